// Interior-update variants measured side by side (afx-interior-ab): the
// leapfrog update of every interior cell of an all-air grid, the operation of
// the library's interior tiles, written several ways, plus a plain copy of
// the same bytes (half read, half written) and a stream with a step's byte
// mix (two fields read, one written, no neighbours).  Each variant's output is
// compared bit for bit with variant A.
//   A  col: one column per thread, z-march with four planes of p^n in flight,
//      x and y neighbours from the L1 / L2 (the library's interior_column);
//   B  vec: V cells per lane along x, x neighbours across lanes by shuffles
//      (the library's interior_vec);
//   C  lds: slab tiles staged through LDS: per plane the tile's p^n plus a
//      one-cell x/y halo goes through LDS (double-buffered, one barrier per
//      plane), z neighbours from registers, loads D planes ahead.
//   D  mstep: T steps per pass (temporal blocking): level k = p^{n+k} over the
//      tile grown by T - k cells in x and y, z-marched with a ring of planes
//      per level in LDS; p^n and p^{n-1} read once, p^{n+T-1} and p^{n+T}
//      written.  Checked against T launches of A (values rounded to the
//      storage type between levels, as stored values are).
//   E  dma (FP32): planes of p^n and p^{n-1} brought into LDS by direct LDS
//      loads (global_load_lds_dword, no VGPRs), A planes ahead, one barrier
//      per plane.
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

// ---- D: T steps per pass.  Level k (k = 1..T) covers the tile grown by
// T - k cells in x and y; the z-march keeps a ring of planes per level in
// LDS (four for p^n, three for each intermediate level).  Iteration j loads
// p^n plane j + 1 and computes level k at plane j - k + 1, k = 1..T, with a
// barrier between levels.  Cells off the grid interior hold 0 at levels >= 1
// (the outer planes are never written).
template <int TX, int TY, int T> struct MsGeom {
    static constexpr int W(int k) { return TX + 2 * (T - k); }
    static constexpr int H(int k) { return TY + 2 * (T - k); }
    static constexpr int N(int k) { return W(k) * H(k); }
    static constexpr int ring(int k) { return k == 0 ? 4 : 3; }
    static constexpr int off(int k) { return k == 0 ? 0 : off(k - 1) + ring(k - 1) * N(k - 1); }
    static constexpr int lds() { return off(T - 1) + ring(T - 1) * N(T - 1); }
};

__device__ __forceinline__ int ms_slot(int z, int ring) { return ring == 4 ? (z & 3) : ((z % 3) + 3) % 3; }

template <typename S, int TX, int TY, int T, int NT, int K>
__device__ __forceinline__ void ms_level(float* L, int j, int x0, int y0, int z0, int TZ, const Room& r, float l2,
                                         float c2, const float* pmc, S* o1, S* o2, long sy, long sz)
{
    using M = MsGeom<TX, TY, T>;
    constexpr int Nk = M::N(K), Wk = M::W(K), Mk = (Nk + NT - 1) / NT, Wp = M::W(K - 1);
    const int z = j - K + 1;
    if (z < z0 - (T - K) || z > z0 + TZ - 1 + (T - K)) return;
    const bool zc = z >= 1 && z < r.nz - 1;
    const float* A = L + M::off(K - 1);
    constexpr int ra = M::ring(K - 1);
    const float* Pc = A + ms_slot(z, ra) * M::N(K - 1);
    const float* Pu = A + ms_slot(z + 1, ra) * M::N(K - 1);
    const float* Pd = A + ms_slot(z - 1, ra) * M::N(K - 1);
#pragma unroll
    for (int m = 0; m < Mk; ++m) {
        const int c = (int)threadIdx.x + m * NT;
        if (c < Nk) {
            const int lx = c % Wk, ly = c / Wk;
            const int x = x0 - (T - K) + lx, y = y0 - (T - K) + ly;
            const int e = (ly + 1) * Wp + lx + 1;
            float v = 0.0f;
            const bool act = zc && x >= 1 && x < r.nx - 1 && y >= 1 && y < r.ny - 1;
            if (act) {
                float s = Pc[e + 1] + Pc[e - 1];
                s = s + Pc[e + Wp];
                s = s + Pc[e - Wp];
                s = s + Pu[e];
                s = s + Pd[e];
                float pm;
                if constexpr (K == 1)
                    pm = pmc[m];
                else
                    pm = (L + M::off(K - 2) + ms_slot(z, M::ring(K - 2)) * M::N(K - 2))[(ly + 2) * M::W(K - 2) + lx + 2];
                v = (float)(S)((l2 * s + c2 * Pc[e]) - pm);
            }
            if constexpr (K < T) {
                (L + M::off(K) + ms_slot(z, 3) * Nk)[c] = v;
            } else if (act) {
                const long gi = (long)z * sz + (long)y * sy + x;
                o2[gi] = (S)v;
                o1[gi] = (S)Pc[e];
            }
        }
    }
}

template <typename S, int TX, int TY, int TZ, int T, int NT, int D>
__global__ __launch_bounds__(NT) void k_mstep(S* __restrict__ o1, S* __restrict__ o2, const S* __restrict__ cur,
                                              const S* __restrict__ prev, Room r, float l2, float c2, int ntx, int nty,
                                              int swz, int ntiles)
{
    using M = MsGeom<TX, TY, T>;
    __shared__ float L[M::lds()];
    const int b = logical_tile(swz);
    if (b >= ntiles) return;
    const int tx = b % ntx, ty = (b / ntx) % nty, tz = b / (ntx * nty);
    const int x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
    const long sy = r.nx, sz = (long)r.nx * r.ny;
    const int tid = (int)threadIdx.x;
    constexpr int M0 = (M::N(0) + NT - 1) / NT, M1 = (M::N(1) + NT - 1) / NT;
    long g0[M0], g1[M1];
    int s0[M0];
    bool v0[M0], a1[M1];
#pragma unroll
    for (int m = 0; m < M0; ++m) {  // p^n over the grown tile (off-grid cells read as 0)
        const int c = tid + m * NT;
        const int cc = c < M::N(0) ? c : M::N(0) - 1;
        const int x = x0 - T + cc % M::W(0), y = y0 - T + cc / M::W(0);
        v0[m] = c < M::N(0) && x >= 0 && x < r.nx && y >= 0 && y < r.ny;
        s0[m] = c < M::N(0) ? c : -1;
        g0[m] = v0[m] ? (long)y * sy + x : 0;
    }
#pragma unroll
    for (int m = 0; m < M1; ++m) {  // p^{n-1} of level 1's cells
        const int c = tid + m * NT;
        const int cc = c < M::N(1) ? c : M::N(1) - 1;
        const int x = x0 - (T - 1) + cc % M::W(1), y = y0 - (T - 1) + cc / M::W(1);
        a1[m] = c < M::N(1) && x >= 1 && x < r.nx - 1 && y >= 1 && y < r.ny - 1;
        g1[m] = a1[m] ? (long)y * sy + x : 0;
    }
    const int jb = z0 - T + 1, je = z0 + TZ + T - 2;
    float ld0[D][M0], pm1[D][M1];
    auto load0 = [&](float (&dst)[M0], int z) {
        z = z < je + 1 ? z : je + 1;
        const bool zv = z >= 0 && z < r.nz;
        const long zo = (long)(zv ? z : 0) * sz;
#pragma unroll
        for (int m = 0; m < M0; ++m) dst[m] = (zv && v0[m]) ? (float)cur[zo + g0[m]] : 0.0f;
    };
    auto store0 = [&](const float (&src)[M0], int z) {
        float* P = L + M::off(0) + (z & 3) * M::N(0);
#pragma unroll
        for (int m = 0; m < M0; ++m)
            if (s0[m] >= 0) P[s0[m]] = src[m];
    };
    auto loadm = [&](float (&dst)[M1], int z) {
        z = z < je ? z : je;
        const bool zv = z >= 1 && z < r.nz - 1;
        const long zo = (long)(zv ? z : 1) * sz;
#pragma unroll
        for (int m = 0; m < M1; ++m) dst[m] = (zv && a1[m]) ? (float)prev[zo + g1[m]] : 0.0f;
    };
    {
        float t0[M0];
        load0(t0, jb - 1);
        store0(t0, jb - 1);
        load0(t0, jb);
        store0(t0, jb);
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
        load0(ld0[(d + 1) % D], jb + 1 + d);
        loadm(pm1[d], jb + d);
    }
    for (int j = jb; j <= je; j += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int jj = j + d;
            if (jj <= je) {
                store0(ld0[(d + 1) % D], jj + 1);
                load0(ld0[(d + 1) % D], jj + 1 + D);
                float pmc[M1];
#pragma unroll
                for (int m = 0; m < M1; ++m) pmc[m] = pm1[d][m];
                loadm(pm1[d], jj + D);
                __syncthreads();
                ms_level<S, TX, TY, T, NT, 1>(L, jj, x0, y0, z0, TZ, r, l2, c2, pmc, o1, o2, sy, sz);
                if constexpr (T >= 2) {
                    __syncthreads();
                    ms_level<S, TX, TY, T, NT, 2>(L, jj, x0, y0, z0, TZ, r, l2, c2, pmc, o1, o2, sy, sz);
                }
                if constexpr (T >= 3) {
                    __syncthreads();
                    ms_level<S, TX, TY, T, NT, 3>(L, jj, x0, y0, z0, TZ, r, l2, c2, pmc, o1, o2, sy, sz);
                }
                if constexpr (T >= 4) {
                    __syncthreads();
                    ms_level<S, TX, TY, T, NT, 4>(L, jj, x0, y0, z0, TZ, r, l2, c2, pmc, o1, o2, sy, sz);
                }
            }
        }
    }
}

// ---- E: one wave per row of a 64 x 4 tile; each plane slot in LDS holds p^n
// rows y0-1 .. y0+4, p^{n-1} rows y0 .. y0+3 and the x halo (left, right),
// 12 direct LDS loads per plane, three per wave, issued A planes ahead into a
// ring of A + 4 slots (a slot is refilled two iterations after its last read,
// across a barrier).  Before computing plane z a wave waits until at most 3A
// of its memory operations are outstanding: every load of plane z + 1 was
// followed by at least 3A more (three loads per plane, stores on top), and
// the counter retires them in order.
#define AB_VM_WAIT(n) __builtin_amdgcn_s_waitcnt(0x0F70 | ((n) & 15) | ((((n) >> 4) & 3) << 14))

template <int TZ, int A>
__global__ __launch_bounds__(256) void k_dma(float* __restrict__ out, const float* __restrict__ cur,
                                             const float* __restrict__ prev, Room r, float l2, float c2, int ntx, int nty,
                                             int swz, int ntiles)
{
    constexpr int TY = 4, R = A + 4;
    constexpr int PS = (2 * TY + 2) * 64 + 2 * 64;  // floats per slot (halo rows padded)
    static_assert(3 * A < 64, "vmcnt field");
    __shared__ float L[R * PS];
    const int b = logical_tile(swz);
    if (b >= ntiles) return;
    const int tx = b % ntx, ty = (b / ntx) % nty, tz = b / (ntx * nty);
    const int x0 = tx * 64, y0 = ty * TY, z0 = tz * TZ;
    const int zs = z0 < 1 ? 1 : z0;
    const int ze = (z0 + TZ) < (r.nz - 1) ? (z0 + TZ) : (r.nz - 1);
    if (zs >= ze) return;
    const long sy = r.nx, sz = (long)r.nx * r.ny;
    const int tid = (int)threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int xl = x0 + lane < r.nx ? x0 + lane : r.nx - 1;
    const float* src[3];
    long off[3];
    int lds[3];
    bool act[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {  // item w + 4k of: p^n rows 0-5, p^{n-1} rows 0-3, halo left, halo right
        const int it = w + 4 * k;
        int yy, xx, lo;
        bool q = false, a = true;
        if (it < TY + 2) {
            yy = y0 - 1 + it; xx = xl; lo = it * 64;
        } else if (it < 2 * TY + 2) {
            yy = y0 + it - (TY + 2); xx = xl; lo = it * 64; q = true;
        } else if (it == 2 * TY + 2) {
            yy = y0 + lane; xx = x0 - 1; lo = (2 * TY + 2) * 64; a = lane < TY;
        } else {
            yy = y0 + lane; xx = x0 + 64; lo = (2 * TY + 3) * 64; a = lane < TY;
        }
        yy = yy < 0 ? 0 : (yy > r.ny - 1 ? r.ny - 1 : yy);
        xx = xx < 0 ? 0 : (xx > r.nx - 1 ? r.nx - 1 : xx);
        src[k] = q ? prev : cur;
        off[k] = (long)yy * sy + xx;
        lds[k] = lo;
        act[k] = a;
    }
    auto issue = [&](int p) {
        const int pc = p < 0 ? 0 : (p > ze ? ze : p);  // past the tile: its last plane again
        const long zo = (long)pc * sz;
        float* slot = L + (p % R) * PS;
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (act[k]) __builtin_amdgcn_global_load_lds(src[k] + zo + off[k], slot + lds[k], 4, 0, 0);
    };
    for (int p = zs - 1; p <= zs + A; ++p) issue(p);
    const int x = x0 + lane, y = y0 + w;
    const bool st_ok = x >= 1 && x <= r.nx - 2 && y >= 1 && y <= r.ny - 2;
    for (int z = zs; z < ze; ++z) {
        issue(z + 1 + A);
        AB_VM_WAIT(3 * A);
        __builtin_amdgcn_s_barrier();
        const float* S0 = L + (z % R) * PS;
        const float* Su = L + ((z + 1) % R) * PS;
        const float* Sd = L + ((z - 1) % R) * PS;
        const int e = (w + 1) * 64 + lane;
        const float c = S0[e];
        const float xp = lane < 63 ? S0[e + 1] : S0[(2 * TY + 3) * 64 + w];
        const float xm = lane > 0 ? S0[e - 1] : S0[(2 * TY + 2) * 64 + w];
        float s = xp + xm;
        s = s + S0[e + 64];
        s = s + S0[e - 64];
        s = s + Su[e];
        s = s + Sd[e];
        const float pm = S0[(TY + 2 + w) * 64 + lane];
        if (st_ok) out[(long)z * sz + (long)y * sy + x] = (l2 * s + c2 * c) - pm;
    }
}

__global__ void k_copy(const float4* __restrict__ a, float4* __restrict__ b, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}

// The step's own byte mix with no stencil: two fields read, one written, 16 B
// per lane (out = a - b, bitwise-meaningless for FP16 but the same bytes).
__global__ void k_stream(const float4* __restrict__ a, const float4* __restrict__ b, float4* __restrict__ c, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float4 u = a[i], v = b[i];
        c[i] = make_float4(u.x - v.x, u.y - v.y, u.z - v.z, u.w - v.w);
    }
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
        fflush(stdout);
    };
    // per-step time of a launch that advances `steps` steps; `ok`: bit-exact
    auto timeit_per = [&](const char* name, auto launch, int steps, bool ok) {
        CK(hipEventRecord(e0));
        for (int k = 0; k < reps; ++k) launch();
        CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
        float ms; CK(hipEventElapsedTime(&ms, e0, e1));
        const double us = ms * 1e3 / reps / steps;
        printf("%s %dx%dx%d %-26s %8.2f us  %7.1f Gcell/s  %6.3f frac  %s (per step)\n", tag, r.nx, r.ny, r.nz, name, us, n / us * 1e-3, 3.0 * sizeof(S) * n / (us * 1e-6) / 8e12, ok ? "bitexact" : "MISMATCH");
        fflush(stdout);
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
        // D: references p^{n+1} .. p^{n+4} from launches of A (XCD order)
        constexpr int TX = 64, TY = 4, TZ = 16;
        const int ntx = (r.nx + TX - 1) / TX, nty = (r.ny + TY - 1) / TY, ntz = (r.nz + TZ - 1) / TZ;
        const int nt = ntx * nty * ntz, ng = (nt + 7) / 8 * 8;
        S* f[4];
        for (auto& p : f) {
            CK(hipMalloc(&p, n * sizeof(S)));
            CK(hipMemset(p, 0, n * sizeof(S)));
        }
        auto col = [&](S* o, const S* c, const S* p) {
            hipLaunchKernelGGL((k_col<S, TX, TY, TZ>), dim3(ng), dim3(TX * TY), 0, 0, o, c, p, r, l2, c2, ntx, nty, 1, nt);
        };
        col(f[0], cur, prev);
        col(f[1], f[0], cur);
        col(f[2], f[1], f[0]);
        col(f[3], f[2], f[1]);
        CK(hipDeviceSynchronize());
        std::vector<std::vector<S>> ref(4, std::vector<S>(n));
        for (int k = 0; k < 4; ++k) CK(hipMemcpy(ref[(size_t)k].data(), f[k], n * sizeof(S), hipMemcpyDeviceToHost));
        for (auto p : f) CK(hipFree(p));
        S *o1, *o2;
        CK(hipMalloc(&o1, n * sizeof(S)));
        CK(hipMalloc(&o2, n * sizeof(S)));
        std::vector<S> g1(n), g2(n);
#define MSTEP(TX_, TY_, TZ_, T_, NT_, D_) { \
        const int mx = (r.nx + TX_ - 1) / TX_, my = (r.ny + TY_ - 1) / TY_, mz = (r.nz + TZ_ - 1) / TZ_; \
        const int mt = mx * my * mz, mg = (mt + 7) / 8 * 8; \
        CK(hipMemset(o1, 0, n * sizeof(S))); CK(hipMemset(o2, 0, n * sizeof(S))); \
        auto L = [&]() { hipLaunchKernelGGL((k_mstep<S, TX_, TY_, TZ_, T_, NT_, D_>), dim3(mg), dim3(NT_), 0, 0, o1, o2, cur, prev, r, l2, c2, mx, my, 1, mt); }; \
        L(); CK(hipDeviceSynchronize()); \
        CK(hipMemcpy(g1.data(), o1, n * sizeof(S), hipMemcpyDeviceToHost)); CK(hipMemcpy(g2.data(), o2, n * sizeof(S), hipMemcpyDeviceToHost)); \
        const bool ok = memcmp(g1.data(), ref[T_ - 2].data(), n * sizeof(S)) == 0 && memcmp(g2.data(), ref[T_ - 1].data(), n * sizeof(S)) == 0; \
        char nm[64]; snprintf(nm, 64, "mstep T%d %dx%dx%d nt%d D%d", T_, TX_, TY_, TZ_, NT_, D_); \
        timeit_per(nm, L, T_, ok); }
        MSTEP(64, 8, 8, 2, 256, 2) MSTEP(64, 8, 16, 2, 256, 2) MSTEP(64, 16, 16, 2, 512, 2) MSTEP(128, 8, 16, 2, 512, 2)
        MSTEP(64, 8, 16, 3, 256, 2) MSTEP(64, 16, 16, 3, 512, 2) MSTEP(64, 16, 16, 4, 512, 2) MSTEP(64, 16, 32, 4, 512, 2)
        MSTEP(64, 16, 32, 4, 512, 4)
        CK(hipFree(o1)); CK(hipFree(o2));
    }
    if constexpr (sizeof(S) == 4) {
#define DMA(TZ_, A_) { \
        const int dx_ = (r.nx + 63) / 64, dy_ = (r.ny + 3) / 4, dz_ = (r.nz + TZ_ - 1) / TZ_; \
        const int dt_ = dx_ * dy_ * dz_, dg_ = (dt_ + 7) / 8 * 8; \
        CK(hipMemset(outB, 0, n * sizeof(S))); \
        auto L = [&]() { hipLaunchKernelGGL((k_dma<TZ_, A_>), dim3(dg_), dim3(256), 0, 0, (float*)outB, (const float*)cur, (const float*)prev, r, l2, c2, dx_, dy_, 1, dt_); }; \
        char nm[64]; snprintf(nm, 64, "dma 64x4x%d A%d xcd", TZ_, A_); timeit(nm, L, outB, true); }
        DMA(16, 2) DMA(16, 4) DMA(16, 6) DMA(32, 4) DMA(8, 4)
    }
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
    {
        // p^n and p^{n-1} read, p^{n+1} written, no neighbours: the byte mix of a step
        const long n16 = sizeof(S) * n / 16;
        auto L = [&]() { hipLaunchKernelGGL(k_stream, dim3(4096), dim3(256), 0, 0, (const float4*)cur, (const float4*)prev, (float4*)outB, n16); };
        timeit("stream 2 read 1 write", L, outB, false);
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
