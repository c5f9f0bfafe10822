/*
 * CPU oracle — TEST INFRASTRUCTURE ONLY (see orc.h).  Built by oracle/build.sh
 * into oracle/liborc.so; loaded only by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg, never by the product library.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "orc.h"

/* Decimal solid-neighbour patterns of lrs-data.hpp:24-51 mapped to the group
 * ids of lrs-data.hpp:53-78 (faces 0-5, edges 6-17, corners 18-25). */
static int orc_group_of(int pattern)
{
    static const int pat[26] = {
        100000, 10000, 1000, 100, 10, 1,
        101000, 100100, 100010, 100001, 11000, 10100, 10010, 10001, 1010, 1001, 110, 101,
        101010, 101001, 100110, 100101, 11010, 11001, 10110, 10101,
    };
    for (int g = 0; g < 26; ++g)
        if (pat[g] == pattern) return g;
    return -1;
}

/* lrs-data.hpp:117-215: scan 1..n-2 on every axis, z outermost; AIR cells
 * with solid neighbours are grouped.  Materials: faces take materials[group],
 * edges/corners take the non-air neighbours in X_P,X_M,Y_P,Y_M,Z_P,Z_M order. */
int64_t orc_classify(int32_t nx, int32_t ny, int32_t nz, const int32_t* v,
                     int32_t* pos, int32_t* group, int64_t cap)
{
    return orc_classify_range(nx, ny, nz, v, nx, ny, nz, pos, group, cap);
}

/* The scan of lrs-data.hpp:154-158 runs over 1 <= x < lim_x - 1 (likewise y,
 * z), where lim is the voxel volume's size (lrs-data.hpp:101-105): a grid
 * grown beyond the volume by minsize keeps plain leapfrog cells there. */
int64_t orc_classify_range(int32_t nx, int32_t ny, int32_t nz, const int32_t* v,
                           int32_t lx, int32_t ly, int32_t lz,
                           int32_t* pos, int32_t* group, int64_t cap)
{
    const int64_t sx = nx, sxy = (int64_t)nx * ny;
    if (lx > nx) lx = nx;
    if (ly > ny) ly = ny;
    if (lz > nz) lz = nz;
    int64_t count[26] = {0};
    /* pass 1: counts per group (and the unphysical check) */
    for (int64_t z = 1; z < lz - 1; ++z)
        for (int64_t y = 1; y < ly - 1; ++y)
            for (int64_t x = 1; x < lx - 1; ++x) {
                const int64_t i = z * sxy + y * sx + x;
                if (v[i] != -1) continue;
                int pattern = 0;
                if (v[i + 1] != -1) pattern += 100000;
                if (v[i - 1] != -1) pattern += 10000;
                if (v[i + sx] != -1) pattern += 1000;
                if (v[i - sx] != -1) pattern += 100;
                if (v[i + sxy] != -1) pattern += 10;
                if (v[i - sxy] != -1) pattern += 1;
                if (pattern == 0) continue;
                int g = orc_group_of(pattern);
                if (g < 0) {
                    if (cap > 0) { pos[0] = (int32_t)x; pos[1] = (int32_t)y; pos[2] = (int32_t)z; }
                    return -1;
                }
                count[g]++;
            }
    int64_t off[27];
    off[0] = 0;
    for (int g = 0; g < 26; ++g) off[g + 1] = off[g] + count[g];
    if (off[26] > cap) return off[26];
    int64_t fill[26];
    for (int g = 0; g < 26; ++g) fill[g] = off[g];
    for (int64_t z = 1; z < lz - 1; ++z)
        for (int64_t y = 1; y < ly - 1; ++y)
            for (int64_t x = 1; x < lx - 1; ++x) {
                const int64_t i = z * sxy + y * sx + x;
                if (v[i] != -1) continue;
                int mats[6] = {-1, -1, -1, -1, -1, -1};
                int pattern = 0;
                if (v[i + 1] != -1) { mats[0] = v[i + 1]; pattern += 100000; }
                if (v[i - 1] != -1) { mats[1] = v[i - 1]; pattern += 10000; }
                if (v[i + sx] != -1) { mats[2] = v[i + sx]; pattern += 1000; }
                if (v[i - sx] != -1) { mats[3] = v[i - sx]; pattern += 100; }
                if (v[i + sxy] != -1) { mats[4] = v[i + sxy]; pattern += 10; }
                if (v[i - sxy] != -1) { mats[5] = v[i - sxy]; pattern += 1; }
                if (pattern == 0) continue;
                int g = orc_group_of(pattern);
                int64_t e = fill[g]++;
                int32_t* q = pos + 6 * e;
                q[0] = (int32_t)x; q[1] = (int32_t)y; q[2] = (int32_t)z;
                q[3] = q[4] = q[5] = -1;
                if (g <= 5) {
                    q[3] = mats[g];
                } else {
                    int k = 0;
                    for (int f = 0; f < 6; ++f)
                        if (mats[f] != -1) q[3 + k++] = mats[f];
                }
                group[e] = g;
            }
    return off[26];
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* Round to nearest, ties to even, into binary16 and back to float.
 * Normal range (|x| >= 2^-14): the 13 significand bits binary16 lacks are
 * rounded off by adding 0x0FFF plus the lowest kept bit and truncating, so
 * a carry out of the significand moves into the exponent; results of 65520
 * and above (beyond the largest finite 65504) become infinity.  Subnormal
 * range: the quantum is 2^-24, the ulp of 0.75f, so (|x| + 0.75f) - 0.75f
 * rounds |x| to it under the default rounding mode. */
static float orc_round_f16(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    const uint32_t sign = u & 0x80000000u, a = u & 0x7FFFFFFFu;
    if (a >= 0x7F800000u) return x; /* inf, nan */
    uint32_t r;
    if (a >= 0x38800000u) {
        r = (a + 0x0FFFu + ((a >> 13) & 1u)) & ~0x1FFFu;
        if (r >= 0x47800000u) r = 0x7F800000u;
    } else {
        float m;
        memcpy(&m, &a, 4);
        volatile float t = m + 0.75f;
        m = t - 0.75f;
        memcpy(&r, &m, 4);
    }
    r |= sign;
    float y;
    memcpy(&y, &r, 4);
    return y;
}

/* Round to nearest, ties to even, into bfloat16 (the upper half of a float)
 * and back: the same carry rounding on the low 16 bits; a carry out of the
 * largest finite value gives infinity.  NaN is returned unchanged. */
static float orc_round_bf16(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return x;
    u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
    float y;
    memcpy(&y, &u, 4);
    return y;
}

float orc_round_store(float x, int32_t mode)
{
    return mode == 1 ? orc_round_f16(x) : (mode == 2 ? orc_round_bf16(x) : x);
}

#define REAL float
#define SFX _f32
#include "orc_impl.h"
#undef REAL
#undef SFX

#define REAL double
#define SFX _f64
#include "orc_impl.h"
#undef REAL
#undef SFX
