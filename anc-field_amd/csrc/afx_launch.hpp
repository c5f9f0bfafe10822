// Host-side check of a kernel launch's geometry, made before every launch of
// the library's kernels: each grid and block dimension positive, at most
// max_threads threads per workgroup (the kernel's launch bound), grid y and
// z within 65535, and the work-item count of every axis within the 32-bit
// grid size of an AQL dispatch packet.  A launch that fails it returns
// hipErrorInvalidConfiguration instead of reaching the queue.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace afx {

inline hipError_t launch_ok(dim3 g, dim3 b, unsigned max_threads)
{
    const uint64_t threads = (uint64_t)b.x * b.y * b.z;
    if (g.x == 0 || g.y == 0 || g.z == 0 || b.x == 0 || b.y == 0 || b.z == 0) return hipErrorInvalidConfiguration;
    if (threads > max_threads || threads > 1024) return hipErrorInvalidConfiguration;
    if (g.y > 65535 || g.z > 65535) return hipErrorInvalidConfiguration;
    if ((uint64_t)g.x * b.x > 0xFFFFFFFFull || (uint64_t)g.y * b.y > 0xFFFFFFFFull ||
        (uint64_t)g.z * b.z > 0xFFFFFFFFull)
        return hipErrorInvalidConfiguration;
    return hipSuccess;
}

}  // namespace afx

#define AFX_CHECK_LAUNCH(g, b, max_threads)                                    \
    do {                                                                       \
        const hipError_t le_ = afx::launch_ok((g), (b), (max_threads));        \
        if (le_ != hipSuccess) return le_;                                     \
    } while (0)
