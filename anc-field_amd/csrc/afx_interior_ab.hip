// Interior-update variants measured side by side (afx-interior-ab): the
// leapfrog update of every interior cell of an all-air grid, the operation of
// the library's interior tiles, written three ways, plus a plain copy of the
// same bytes.  Each variant's output is compared bit for bit with variant A.
//   A  col: one column per thread, z-march with four planes of p^n in flight,
//      x and y neighbours from the L1 / L2 (the library's interior_column);
//   B  vec: V cells per lane along x, x neighbours across lanes by shuffles
//      (the library's interior_vec);
//   C  lds: slab tiles staged through LDS: per plane the tile's p^n plus a
//      one-cell x/y halo goes through LDS (double-buffered, one barrier per
//      plane), z neighbours from registers, loads D planes ahead.
// Grids: 256x192x128 and 640x480x320, FP32 and FP16 storage (FP32 compute).
// Timing: HIP events around 200 back-to-back launches after one warm-up.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <cstring>
#include <cmath>

#define CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { printf("HIP %s at %d\n", hipGetErrorString(e_), __LINE__); exit(1);} } while (0)

struct Room { int nx, ny, nz; };

// ---- A: scalar column (4 planes in flight)
// XCD-aware order: block b runs on XCD b % 8; give each XCD a contiguous
// run of logical tiles (x fastest, then y, then z) so y/z halos share L2.
__device__ __forceinline__ int logical_tile(int swz)
{
    const int b = blockIdx.x, nb = gridDim.x;
    if (!swz) return b;
    const int per = (nb + 7) / 8;
    return (b % 8) * per + b / 8;
}

template <typename S, int TX, int TY, int TZ>
__global__ __launch_bounds__(TX * TY) void k_col(S* __restrict__ out, const S* __restrict__ cur,
                                                 const S* __restrict__ prev, Room r, float l2, float c2, int ntx, int nty, int swz, int ntiles)
{
    const int b = logical_tile(swz);
    if (b >= ntiles) return;
    const int tx = b % ntx, ty = (b / ntx) % nty, tz = b / (ntx * nty);
    const int x = tx * TX + (int)(threadIdx.x % TX), y = ty * TY + (int)(threadIdx.x / TX), z0 = tz * TZ;
    if (x < 1 || x >= r.nx - 1 || y < 1 || y >= r.ny - 1) return;
    const int zs = z0 < 1 ? 1 : z0;
    const int ze = (z0 + TZ) < (r.nz - 1) ? (z0 + TZ) : (r.nz - 1);
    if (zs >= ze) return;
    const long sy = r.nx, sz = (long)r.nx * r.ny;
    long i = (long)zs * sz + (long)y * sy + x;
    constexpr int kDepth = 4;
    struct Plane { float u, xp, xm, yp, ym, pm; };
    auto load_plane = [&](Plane& P, long j) {
        P.u = (float)cur[j + sz]; P.xp = (float)cur[j + 1]; P.xm = (float)cur[j - 1]; P.yp = (float)cur[j + sy]; P.ym = (float)cur[j - sy]; P.pm = (float)prev[j];
    };
    const long i_last = i + (long)(ze - 1 - zs) * sz;
    float pd = (float)cur[i - sz], pc = (float)cur[i];
    Plane P[kDepth];
#pragma unroll
    for (int d = 0; d < kDepth; ++d) load_plane(P[d], zs + d < ze ? i + (long)d * sz : i_last);
    for (int z = zs; z < ze; z += kDepth) {
#pragma unroll
        for (int d = 0; d < kDepth; ++d) {
            const long j = i + (long)d * sz;
            if (z + d < ze) {
                float s = P[d].xp + P[d].xm; s = s + P[d].yp; s = s + P[d].ym; s = s + P[d].u; s = s + pd;
                out[j] = (S)((l2 * s + c2 * pc) - P[d].pm);
                pd = pc; pc = P[d].u;
            }
            const long jn = z + d + kDepth < ze ? j + (long)kDepth * sz : i_last;
            load_plane(P[d], jn);
        }
        i += (long)kDepth * sz;
    }
}

// ---- B: V cells per lane along x, neighbours across lanes by shuffles
template <typename S, int V> struct Vec { S v[V]; };
template <typename S, int V>
__device__ __forceinline__ void ldv(float* o, const S* p) {
    Vec<S, V> t; __builtin_memcpy(&t, p, sizeof(t));
#pragma unroll
    for (int k = 0; k < V; ++k) o[k] = (float)t.v[k];
}

template <typename S, int V, int TY, int TZ, int DEPTH>
__global__ __launch_bounds__(64 * TY) void k_vec(S* __restrict__ out, const S* __restrict__ cur,
                                                 const S* __restrict__ prev, Room r, float l2, float c2, int ntx, int nty, int swz, int ntiles)
{
    constexpr int TXC = 64 * V;
    const int b = logical_tile(swz);
    if (b >= ntiles) return;
    const int tx = b % ntx, ty = (b / ntx) % nty, tz = b / (ntx * nty);
    const int lane = (int)threadIdx.x & 63;
    const int x = tx * TXC + lane * V;
    const int y = ty * TY + (int)(threadIdx.x >> 6), z0 = tz * TZ;
    if (y < 1 || y >= r.ny - 1) return;  // whole wave
    const int zs = z0 < 1 ? 1 : z0;
    const int ze = (z0 + TZ) < (r.nz - 1) ? (z0 + TZ) : (r.nz - 1);
    if (zs >= ze) return;
    const long sy = r.nx, sz = (long)r.nx * r.ny;
    const int xl = x + V - 1 < r.nx ? x : r.nx - V;
    long i = (long)zs * sz + (long)y * sy + xl;
    const int xe = lane == 0 ? (x - 1 >= 0 ? x - 1 : 0) : (x + V < r.nx ? x + V : r.nx - 1);
    const long ie = (long)zs * sz + (long)y * sy + xe;
    struct Plane { float u[V], yp[V], ym[V], pm[V]; float e; };
    auto load_plane = [&](Plane& P, long j, long je) {
        ldv<S, V>(P.u, cur + j + sz); ldv<S, V>(P.yp, cur + j + sy); ldv<S, V>(P.ym, cur + j - sy); ldv<S, V>(P.pm, prev + j);
        P.e = (float)cur[je + sz];
    };
    const long i_last = i + (long)(ze - 1 - zs) * sz, ie_last = ie + (long)(ze - 1 - zs) * sz;
    float pd[V], pc[V];
    ldv<S, V>(pd, cur + i - sz); ldv<S, V>(pc, cur + i);
    float ec = (float)cur[ie];
    Plane P[DEPTH];
#pragma unroll
    for (int d = 0; d < DEPTH; ++d) load_plane(P[d], zs + d < ze ? i + (long)d * sz : i_last, zs + d < ze ? ie + (long)d * sz : ie_last);
    const bool full = x >= 1 && x + V - 1 <= r.nx - 2;
    for (int z = zs; z < ze; z += DEPTH) {
#pragma unroll
        for (int d = 0; d < DEPTH; ++d) {
            const long j = i + (long)d * sz;
            if (z + d < ze) {
                const float left_in = __shfl_up(pc[V - 1], 1);
                const float right_in = __shfl_down(pc[0], 1);
                const float left = lane == 0 ? ec : left_in;
                const float right = lane == 63 ? ec : right_in;
                Vec<S, V> o;
#pragma unroll
                for (int k = 0; k < V; ++k) {
                    const float xp = k == V - 1 ? right : pc[k + 1];
                    const float xm = k == 0 ? left : pc[k - 1];
                    float s = xp + xm; s = s + P[d].yp[k]; s = s + P[d].ym[k]; s = s + P[d].u[k]; s = s + pd[k];
                    o.v[k] = (S)((l2 * s + c2 * pc[k]) - P[d].pm[k]);
                }
                if (full) {
                    __builtin_memcpy(out + j, &o, sizeof(o));
                } else {
#pragma unroll
                    for (int k = 0; k < V; ++k)
                        if (x + k >= 1 && x + k <= r.nx - 2 && xl == x) out[j + k] = o.v[k];
                }
#pragma unroll
                for (int k = 0; k < V; ++k) { pd[k] = pc[k]; pc[k] = P[d].u[k]; }
                ec = P[d].e;
            }
            const bool more = z + d + DEPTH < ze;
            load_plane(P[d], more ? j + (long)DEPTH * sz : i_last, more ? ie + (long)(z + d + DEPTH - zs) * sz : ie_last);
        }
        i += (long)DEPTH * sz;
    }
}

// ---- C: LDS-staged slab tile: each thread one column; per plane the tile's
// p^n plus a one-cell x/y halo goes through LDS (double-buffered, one barrier
// per plane); z neighbours from registers; loads D planes ahead
template <typename S, int TX, int TY, int TZ, int D>
__global__ __launch_bounds__(TX * TY) void k_lds(S* __restrict__ out, const S* __restrict__ cur,
                                                 const S* __restrict__ prev, Room r, float l2, float c2, int ntx, int nty, int swz, int ntiles)
{
    constexpr int NT = TX * TY, LX = TX + 2, LY = TY + 2, H = 2 * TX + 2 * TY;
    static_assert(H <= NT, "halo");
    __shared__ float L[2][LX * LY];
    const int b = logical_tile(swz);
    if (b >= ntiles) return;
    const int tx = b % ntx, ty = (b / ntx) % nty, tz = b / (ntx * nty);
    const int x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
    const int t = (int)threadIdx.x, lx = t % TX, ly = t / TX;
    const int x = x0 + lx, y = y0 + ly;
    const int zs = z0 < 1 ? 1 : z0;
    const int ze = (z0 + TZ) < (r.nz - 1) ? (z0 + TZ) : (r.nz - 1);
    if (zs >= ze) return;
    const long sy = r.nx, sz = (long)r.nx * r.ny;
    const int xc = x < r.nx - 1 ? x : r.nx - 1, yc = y < r.ny - 1 ? y : r.ny - 1;
    const bool st_ok = x >= 1 && x <= r.nx - 2 && y >= 1 && y <= r.ny - 2;
    const long io = (long)yc * sy + xc;
    const int lo = (ly + 1) * LX + (lx + 1);
    const bool has_h = t < H;
    int hx, hy, hl;
    if (t < TX) { hx = x0 + t; hy = y0 - 1; hl = t + 1; }
    else if (t < 2 * TX) { hx = x0 + t - TX; hy = y0 + TY; hl = (TY + 1) * LX + (t - TX) + 1; }
    else if (t < 2 * TX + TY) { hx = x0 - 1; hy = y0 + (t - 2 * TX); hl = (t - 2 * TX + 1) * LX; }
    else { hx = x0 + TX; hy = y0 + (t - 2 * TX - TY); hl = (t - 2 * TX - TY + 1) * LX + TX + 1; }
    hx = hx < 0 ? 0 : (hx > r.nx - 1 ? r.nx - 1 : hx);
    hy = hy < 0 ? 0 : (hy > r.ny - 1 ? r.ny - 1 : hy);
    const long ih = has_h ? (long)hy * sy + hx : io;
    struct Pl { float c, h, m; };
    auto load = [&](Pl& P, int zp) {
        const long o = (long)zp * sz;
        P.c = (float)cur[o + io]; P.h = (float)cur[o + ih]; P.m = (float)prev[o + io];
    };
    float pd = (float)cur[(long)(zs - 1) * sz + io];
    Pl Q; load(Q, zs);
    float pc = Q.c, hc = Q.h, pmc = Q.m;
    Pl P[D];
#pragma unroll
    for (int d = 0; d < D; ++d) load(P[d], zs + 1 + d <= ze ? zs + 1 + d : ze);
    for (int z = zs; z < ze; z += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int zz = z + d;
            if (zz < ze) {
                float* B = L[zz & 1];
                B[lo] = pc;
                if (has_h) B[hl] = hc;
                __syncthreads();
                float s = B[lo + 1] + B[lo - 1]; s = s + B[lo + LX]; s = s + B[lo - LX]; s = s + P[d].c; s = s + pd;
                if (st_ok) out[(long)zz * sz + io] = (S)((l2 * s + c2 * pc) - pmc);
                pd = pc; pc = P[d].c; hc = P[d].h; pmc = P[d].m;
                const int zn = zz + 1 + D <= ze ? zz + 1 + D : ze;
                load(P[d], zn);
            }
        }
    }
}

__global__ void k_copy(const float4* __restrict__ a, float4* __restrict__ b, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}

template <typename S>
void run(Room r, const char* tag)
{
    const long n = (long)r.nx * r.ny * r.nz;
    std::vector<S> h(n);
    srand(1);
    S *cur, *prev, *outA, *outB;
    CK(hipMalloc(&cur, n * sizeof(S))); CK(hipMalloc(&prev, n * sizeof(S))); CK(hipMalloc(&outA, n * sizeof(S))); CK(hipMalloc(&outB, n * sizeof(S)));
    for (long i = 0; i < n; ++i) h[i] = (S)((float)rand() / (float)RAND_MAX - 0.5f);
    CK(hipMemcpy(cur, h.data(), n * sizeof(S), hipMemcpyHostToDevice));
    for (long i = 0; i < n; ++i) h[i] = (S)((float)rand() / (float)RAND_MAX - 0.5f);
    CK(hipMemcpy(prev, h.data(), n * sizeof(S), hipMemcpyHostToDevice));
    CK(hipMemset(outA, 0, n * sizeof(S))); CK(hipMemset(outB, 0, n * sizeof(S)));
    const float l2 = 0.32667f, c2 = 2.0f * (1.0f - 3.0f * l2);
    hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
    const int reps = 200;
    std::vector<S> ha(n), hb(n);
    auto timeit = [&](const char* name, auto launch, void* out, bool check) {
        launch(); CK(hipDeviceSynchronize());
        CK(hipEventRecord(e0));
        for (int k = 0; k < reps; ++k) launch();
        CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
        float ms; CK(hipEventElapsedTime(&ms, e0, e1));
        const double us = ms * 1e3 / reps;
        const char* ok = "";
        if (check) {
            CK(hipMemcpy(hb.data(), out, n * sizeof(S), hipMemcpyDeviceToHost));
            ok = memcmp(ha.data(), hb.data(), n * sizeof(S)) == 0 ? "bitexact" : "MISMATCH";
        }
        printf("%s %dx%dx%d %-26s %8.2f us  %7.1f Gcell/s  %6.3f frac  %s\n", tag, r.nx, r.ny, r.nz, name, us, n / us * 1e-3, 3.0 * sizeof(S) * n / (us * 1e-6) / 8e12, ok);
    };
    // A: reference
    for (int swz = 0; swz < 2; ++swz) {
        constexpr int TX = 64, TY = 4, TZ = 16;
        const int ntx = (r.nx + TX - 1) / TX, nty = (r.ny + TY - 1) / TY, ntz = (r.nz + TZ - 1) / TZ;
        const int nt = ntx * nty * ntz, ng = (nt + 7) / 8 * 8;
        auto L = [&]() { hipLaunchKernelGGL((k_col<S, TX, TY, TZ>), dim3(ng), dim3(TX * TY), 0, 0, outA, cur, prev, r, l2, c2, ntx, nty, swz, nt); };
        timeit(swz ? "col 64x4x16 xcd" : "col 64x4x16", L, outA, false);
        CK(hipMemcpy(ha.data(), outA, n * sizeof(S), hipMemcpyDeviceToHost));
    }
#define VEC(V, TY, TZ, D) for (int swz = 0; swz < 2; ++swz) { \
        constexpr int TXC = 64 * V; \
        const int ntx = (r.nx + TXC - 1) / TXC, nty = (r.ny + TY - 1) / TY, ntz = (r.nz + TZ - 1) / TZ; \
        const int nt = ntx * nty * ntz, ng = (nt + 7) / 8 * 8; \
        CK(hipMemset(outB, 0, n * sizeof(S))); \
        auto L = [&]() { hipLaunchKernelGGL((k_vec<S, V, TY, TZ, D>), dim3(ng), dim3(64 * TY), 0, 0, outB, cur, prev, r, l2, c2, ntx, nty, swz, nt); }; \
        char nm[64]; snprintf(nm, 64, "vec V%d TY%d TZ%d D%d%s", V, TY, TZ, D, swz ? " xcd" : ""); timeit(nm, L, outB, true); }
#define LDS(TX, TY, TZ, D) for (int swz = 1; swz < 2; ++swz) {         const int ntx = (r.nx + TX - 1) / TX, nty = (r.ny + TY - 1) / TY, ntz = (r.nz + TZ - 1) / TZ;         const int nt = ntx * nty * ntz, ng = (nt + 7) / 8 * 8;         CK(hipMemset(outB, 0, n * sizeof(S)));         auto L = [&]() { hipLaunchKernelGGL((k_lds<S, TX, TY, TZ, D>), dim3(ng), dim3(TX * TY), 0, 0, outB, cur, prev, r, l2, c2, ntx, nty, swz, nt); };         char nm[64]; snprintf(nm, 64, "lds %dx%dx%d D%d%s", TX, TY, TZ, D, swz ? " xcd" : ""); timeit(nm, L, outB, true); }
    LDS(64, 4, 16, 2) LDS(64, 4, 16, 4) LDS(64, 8, 16, 2) LDS(64, 8, 32, 2) LDS(128, 4, 16, 2) LDS(64, 4, 32, 4) LDS(32, 8, 16, 2) LDS(128, 8, 16, 2)
    VEC(4, 4, 8, 4) VEC(4, 4, 8, 2) VEC(4, 2, 8, 2) VEC(4, 4, 4, 2) VEC(4, 4, 16, 2) VEC(2, 4, 8, 2) VEC(2, 4, 16, 2) VEC(8, 4, 8, 2) VEC(8, 2, 8, 2)
    {
        // read 6n bytes, write 6n bytes: the same 12 B per cell
        const long n16 = 3 * sizeof(S) * n / 2 / 16;
        float *src, *dst;
        CK(hipMalloc(&src, n16 * 16)); CK(hipMalloc(&dst, n16 * 16));
        CK(hipMemset(src, 0, n16 * 16));
        auto L = [&]() { hipLaunchKernelGGL(k_copy, dim3(8192), dim3(256), 0, 0, (const float4*)src, (float4*)dst, n16); };
        timeit("copy (same bytes)", L, dst, false);
        CK(hipFree(src)); CK(hipFree(dst));
    }
    CK(hipFree(cur)); CK(hipFree(prev)); CK(hipFree(outA)); CK(hipFree(outB));
}

int main(int argc, char** argv)
{
    run<float>(Room{256, 192, 128}, "f32");
    run<_Float16>(Room{256, 192, 128}, "f16");
    run<float>(Room{640, 480, 320}, "f32");
    run<_Float16>(Room{640, 480, 320}, "f16");
    return 0;
}
