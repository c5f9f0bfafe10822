// gfx950 kernels of the FDTD room-acoustics step.
//
// Reference counterparts (paths relative to the reference root):
//   k_interior   anc-field/src/leapfrog.cl:34-55
//   bnd_value    anc-field/src/lrs_1.cl:116-179, lrs_2.cl:62-215, lrs_3.cl:62-186
//   special      anc-field/src/source.cl:32-40, mic.cl:32-41 (order simulation.hpp:384-392)
//   k_fused      the whole step of simulation.hpp:358-392 in one launch
//   pml_compat   anc-field/src/pml.cl:154-265
//   pml_phys     no counterpart (Grote-Sim PML restricted to its slabs)
//
// Arithmetic follows the reference's operation order and the library is
// compiled with -ffp-contract=off and IEEE division, so FP32 results are
// bit-identical to the reference operations evaluated in IEEE single.
#include "afx_internal.hpp"
#include "afx_launch.hpp"

namespace afx {

const TileShape kTiles[kNumTiles] = {
    {64, 4, 16, 1}, {64, 4, 32, 1}, {64, 8, 16, 1}, {64, 2, 32, 1}, {128, 2, 16, 1}, {64, 4, 8, 1}, {256, 1, 16, 1},
    {256, 4, 8, 4}, {256, 2, 8, 4}, {128, 4, 16, 2}, {128, 4, 8, 2},
};

// ---------------------------------------------------------------- storage I/O
template <typename S> struct Compute { using T = float; };
template <> struct Compute<double> { using T = double; };

template <typename S, typename C>
__device__ __forceinline__ C ld(const S* __restrict__ p, int64_t i) { return static_cast<C>(p[i]); }
template <typename S, typename C>
__device__ __forceinline__ void st(S* __restrict__ p, int64_t i, C v) { p[i] = static_cast<S>(v); }
// value after a round trip through storage precision (identity for FP32/FP64)
template <typename S, typename C>
__device__ __forceinline__ C rt(C v) { return static_cast<C>(static_cast<S>(v)); }

// ---------------------------------------------------------------- PML value
// Physical PML state (compute type C).  phi_x lives at x+1/2 and is stored
// with cell x (likewise y, z); psi is cell-centred at half steps.  phi and psi
// are compact: only the PML region is stored, box by box (pml_find).
template <typename C> struct PmlDev {
    C* phx;
    C* phy;
    C* phz;
    C* psi;
    const C* z[3];
    C dt, dx;
};

// Grote-Sim pressure update of one PML cell: `sum` is the 6-neighbour sum of
// p^n, pc/pm the cell's p^n/p^{n-1}, dv the divergence of phi^n times dx,
// ps = psi^{n-1/2} and psn = psi^{n+1/2} = ps + dt*pc.
template <typename C>
__device__ __forceinline__ C pml_formula(C sum, C pc, C pm, C dv, C ps, C psn, C sx, C syy, C szz, C dt, C dx, C l2)
{
    const C s1 = sx + syy + szz;
    const C s2 = sx * syy + syy * szz + szz * sx;
    const C s3 = sx * syy * szz;
    C num = l2 * (sum - (C)6.0 * pc);
    num = num + (dt * dt / dx) * dv;
    num = num - (dt * dt) * s3 * (C)0.5 * (psn + ps);
    num = num - (dt * dt) * s2 * pc;
    num = num + (C)2.0 * pc - pm;
    num = num + (C)0.5 * dt * s1 * pm;
    return num / ((C)1.0 + (C)0.5 * dt * s1);
}

// Split path: phi^n and psi read from single buffers, psi advanced in place.
// k is the cell's compact index, kxm/kym/kzm those of its -x/-y/-z
// neighbours (-1 outside the PML region, where phi is zero).
template <typename C>
__device__ __forceinline__ C pml_value(C sum, C pc, C pm, int64_t k, int64_t kxm, int64_t kym, int64_t kzm, C sx,
                                       C syy, C szz, const PmlDev<C>& q, C l2)
{
    const C dv = (q.phx[k] - (kxm >= 0 ? q.phx[kxm] : (C)0)) + (q.phy[k] - (kym >= 0 ? q.phy[kym] : (C)0)) +
                 (q.phz[k] - (kzm >= 0 ? q.phz[kzm] : (C)0));
    const C ps = q.psi[k];
    const C psn = ps + q.dt * pc;
    q.psi[k] = psn;
    return pml_formula<C>(sum, pc, pm, dv, ps, psn, sx, syy, szz, q.dt, q.dx, l2);
}

// Compact index of cell (x, y, z) of room r in the PML phi/psi arrays, or -1
// outside the room's PML region.  boxes: [nbox][8] = room, x0, x1, y0, y1,
// z0, z1 (half-open), base; a box is stored x-fastest from base.  The boxes
// of a room are disjoint, cover its PML region and sit at box0 .. box1 of the
// table; `hint` (a box of the same room, usually the cell's own) is tried first.
__device__ __forceinline__ int64_t pml_find(const int32_t* __restrict__ boxes, const RoomDev& r, int hint, int x,
                                            int y, int z)
{
    for (int k = -1; k < r.box1 - r.box0; ++k) {
        const int32_t* b = boxes + 8 * (k < 0 ? hint : r.box0 + k);
        if (x >= b[1] && x < b[2] && y >= b[3] && y < b[4] && z >= b[5] && z < b[6])
            return (int64_t)b[7] + ((int64_t)(z - b[5]) * (b[4] - b[3]) + (y - b[3])) * (b[2] - b[1]) + (x - b[1]);
    }
    return -1;
}

// Crank-Nicolson step of one phi component at a +1/2 point: `old` is phi^n,
// (p1j, p1n) p^{n+1} at the cell and its + neighbour, (pj, pn) p^n there,
// (psj, psn) psi^{n+1/2} there; h zeta at the half point, g2 and g3 the
// cross-axis factors.
template <typename C>
__device__ __forceinline__ C phi_update(C old, C p1j, C p1n, C pj, C pn, C psj, C psn, C h, C g2, C g3, C cc,
                                        C idt, C idx)
{
    const C dn = (p1n - p1j) * idx, dold = (pn - pj) * idx;
    const C dps = (psn - psj) * idx;
    return ((idt - (C)0.5 * h) * old + cc * g2 * (C)0.5 * (dn + dold) + cc * g3 * dps) / (idt + (C)0.5 * h);
}

// ---------------------------------------------------------------- interior
// One thread owns an (x, y) column of a TX x TY x TZ tile and marches z,
// carrying p^n at z-1, z, z+1 in registers; x and y neighbours come from the
// same and adjacent 64-lane rows (L1 / L2 hits).  The sum order is the
// reference's: x+1, x-1, y+1, y-1, z+1, z-1.  Cells whose bit z - z0 is set
// in `skipbits` (boundary, source/receiver and PML cells, finished by other
// threads of the same launch) are left alone.
template <typename S, typename C, int TZ>
__device__ __forceinline__ void interior_column(S* __restrict__ out, const S* __restrict__ cur,
                                                const S* __restrict__ prev, const RoomDev& r, int x, int y,
                                                int z0, C l2, C c2, uint32_t skipbits, int zlo = 0, int zhi = TZ)
{
    if (x < 1 || x >= r.nx - 1 || y < 1 || y >= r.ny - 1) return;
    const int zs = z0 + zlo < 1 ? 1 : z0 + zlo;
    const int ze = (z0 + zhi) < (r.nz - 1) ? (z0 + zhi) : (r.nz - 1);
    if (zs >= ze) return;
    const int64_t sy = r.nx, sz = (int64_t)r.nx * r.ny;
    int64_t i = r.base + (int64_t)zs * sz + (int64_t)y * sy + x;
    // Software pipeline over a ring of kDepth register sets: the loads of
    // plane z+kDepth are issued right after plane z is computed, so kDepth
    // planes are in flight per thread.  Prefetch addresses clamp to the last
    // plane, so no load sits behind a branch; a plane past the end is not
    // computed (a uniform branch around arithmetic and the store only).
    constexpr int kDepth = 4;
    struct Plane {
        C u, xp, xm, yp, ym, pm;  // p^n at z+1, x+-1, y+-1; p^{n-1}
    };
    auto load_plane = [&](Plane& P, int64_t j) {
        P.u = ld<S, C>(cur, j + sz);
        P.xp = ld<S, C>(cur, j + 1);
        P.xm = ld<S, C>(cur, j - 1);
        P.yp = ld<S, C>(cur, j + sy);
        P.ym = ld<S, C>(cur, j - sy);
        P.pm = ld<S, C>(prev, j);
    };
    const int64_t i_last = i + (int64_t)(ze - 1 - zs) * sz;
    C pd = ld<S, C>(cur, i - sz);  // p^n at z-1 and z of the next plane to compute
    C pc = ld<S, C>(cur, i);
    Plane P[kDepth];
#pragma unroll
    for (int d = 0; d < kDepth; ++d) load_plane(P[d], zs + d < ze ? i + (int64_t)d * sz : i_last);
    for (int z = zs; z < ze; z += kDepth) {
#pragma unroll
        for (int d = 0; d < kDepth; ++d) {
            const int64_t j = i + (int64_t)d * sz;
            if (z + d < ze) {  // uniform across the wave
                C s = P[d].xp + P[d].xm;
                s = s + P[d].yp;
                s = s + P[d].ym;
                s = s + P[d].u;
                s = s + pd;
                const bool skip = (skipbits >> (z + d - z0)) & 1u;
                if (!skip) {
                    const C v = l2 * s + c2 * pc;
                    st<S, C>(out, j, v - P[d].pm);
                }
                pd = pc;
                pc = P[d].u;
            }
            const int64_t jn = z + d + kDepth < ze ? j + (int64_t)kDepth * sz : i_last;
            load_plane(P[d], jn);
        }
        i += (int64_t)kDepth * sz;
    }
}

// V consecutive cells of x per lane, one 64-lane wave per row of the tile:
// the lane's V cells of each plane are read and written as one vector, and
// the x neighbours across lanes come from the adjacent lanes' registers
// (lane 0 and lane 63 read their outer neighbour).  Otherwise as
// interior_column: z-march with p^n at z-1, z, z+1 in registers, planes
// prefetched two ahead, the same operations in the same order per cell, and
// cells whose skip bit is set left alone.  Needs x, nx and the room base to
// be multiples of V (aligned vectors); lanes past the row's end read the
// row's last vector and store nothing.
template <typename S, int V> struct alignas(sizeof(S) * V) VecPack {
    S v[V];
};

template <typename S, typename C, int V>
__device__ __forceinline__ void ldv(C (&o)[V], const S* __restrict__ p, int64_t i)
{
    const VecPack<S, V> t = *reinterpret_cast<const VecPack<S, V>*>(p + i);
#pragma unroll
    for (int k = 0; k < V; ++k) o[k] = static_cast<C>(t.v[k]);
}

template <typename S, typename C, int V, int TZ>
__device__ __forceinline__ void interior_vec(S* __restrict__ out, const S* __restrict__ cur,
                                             const S* __restrict__ prev, const RoomDev& r, int x, int y, int z0,
                                             C l2, C c2, const uint32_t (&skip)[V], int zlo = 0, int zhi = TZ)
{
    if (y < 1 || y >= r.ny - 1) return;  // the whole wave: one row
    const int zs = z0 + zlo < 1 ? 1 : z0 + zlo;
    const int ze = (z0 + zhi) < (r.nz - 1) ? (z0 + zhi) : (r.nz - 1);
    if (zs >= ze) return;
    const int lane = (int)(threadIdx.x & 63u);
    const bool inrow = x < r.nx;
    const int64_t sy = r.nx, sz = (int64_t)r.nx * r.ny;
    int64_t i = r.base + (int64_t)zs * sz + (int64_t)y * sy + (inrow ? x : r.nx - V);
    const int xe = lane == 0 ? (x >= 1 ? x - 1 : 0) : (x + V < r.nx ? x + V : r.nx - 1);
    const int64_t ie = r.base + (int64_t)zs * sz + (int64_t)y * sy + xe;
    bool can[V];  // cell inside the row's updated range [1, nx-2]
    bool all = inrow;
#pragma unroll
    for (int k = 0; k < V; ++k) {
        can[k] = inrow && x + k >= 1 && x + k <= r.nx - 2;
        all = all && can[k];
    }
    constexpr int kDepth = 2;
    struct Plane {
        C u[V], yp[V], ym[V], pm[V];  // p^n at z+1, y+1, y-1; p^{n-1}
        C e;                          // p^n of the lane's outer x neighbour at z+1
    };
    auto load_plane = [&](Plane& P, int64_t j, int64_t je) {
        ldv<S, C, V>(P.u, cur, j + sz);
        ldv<S, C, V>(P.yp, cur, j + sy);
        ldv<S, C, V>(P.ym, cur, j - sy);
        ldv<S, C, V>(P.pm, prev, j);
        P.e = ld<S, C>(cur, je + sz);
    };
    const int64_t i_last = i + (int64_t)(ze - 1 - zs) * sz, ie_last = ie + (int64_t)(ze - 1 - zs) * sz;
    C pd[V], pc[V];
    ldv<S, C, V>(pd, cur, i - sz);
    ldv<S, C, V>(pc, cur, i);
    C ec = ld<S, C>(cur, ie);
    Plane P[kDepth];
#pragma unroll
    for (int d = 0; d < kDepth; ++d)
        load_plane(P[d], zs + d < ze ? i + (int64_t)d * sz : i_last, zs + d < ze ? ie + (int64_t)d * sz : ie_last);
    for (int z = zs; z < ze; z += kDepth) {
#pragma unroll
        for (int d = 0; d < kDepth; ++d) {
            const int64_t j = i + (int64_t)d * sz;
            if (z + d < ze) {  // uniform across the wave
                const C left_in = __shfl_up(pc[V - 1], 1);
                const C right_in = __shfl_down(pc[0], 1);
                const C left = lane == 0 ? ec : left_in;
                const C right = lane == 63 ? ec : right_in;
                VecPack<S, V> o;
                bool full = all;
#pragma unroll
                for (int k = 0; k < V; ++k) {
                    const C xp = k == V - 1 ? right : pc[k + 1];
                    const C xm = k == 0 ? left : pc[k - 1];
                    C s = xp + xm;
                    s = s + P[d].yp[k];
                    s = s + P[d].ym[k];
                    s = s + P[d].u[k];
                    s = s + pd[k];
                    const C v = l2 * s + c2 * pc[k];
                    o.v[k] = static_cast<S>(v - P[d].pm[k]);
                    full = full && !((skip[k] >> (z + d - z0)) & 1u);
                }
                if (full) {
                    *reinterpret_cast<VecPack<S, V>*>(out + j) = o;
                } else {
#pragma unroll
                    for (int k = 0; k < V; ++k)
                        if (can[k] && !((skip[k] >> (z + d - z0)) & 1u)) out[j + k] = o.v[k];
                }
#pragma unroll
                for (int k = 0; k < V; ++k) {
                    pd[k] = pc[k];
                    pc[k] = P[d].u[k];
                }
                ec = P[d].e;
            }
            const bool more = z + d + kDepth < ze;
            load_plane(P[d], more ? j + (int64_t)kDepth * sz : i_last,
                       more ? ie + (int64_t)(z + d + kDepth - zs) * sz : ie_last);
        }
        i += (int64_t)kDepth * sz;
    }
}

// The interior cells of one tile (origin ox, oy, oz of room r), one column
// per thread or V columns per lane; cmask: the tile slot's column masks
// (TX x TY words, bit z - oz set for cells left to a list), or null.  Only
// planes oz + zlo .. oz + zhi - 1 are updated (a tile may be dealt out in
// parts).
template <typename S, typename C, int TX, int TY, int TZ, int V>
__device__ __forceinline__ void tile_interior(S* __restrict__ out, const S* __restrict__ cur,
                                              const S* __restrict__ prev, const RoomDev& r, int ox, int oy, int oz,
                                              C l2, C c2, const uint32_t* __restrict__ cmask, int zlo = 0,
                                              int zhi = TZ)
{
    if constexpr (V == 1) {
        const int x = ox + (int)(threadIdx.x % TX);
        const int y = oy + (int)(threadIdx.x / TX);
        const uint32_t bits = cmask ? cmask[threadIdx.x] : 0u;
        interior_column<S, C, TZ>(out, cur, prev, r, x, y, oz, l2, c2, bits, zlo, zhi);
    } else {
        static_assert(TX == 64 * V, "one 64-lane wave per tile row");
        const int lane = (int)(threadIdx.x & 63u), row = (int)(threadIdx.x >> 6);
        uint32_t bits[V];
        if (cmask) {
            const VecPack<uint32_t, V> w =
                *reinterpret_cast<const VecPack<uint32_t, V>*>(cmask + row * TX + lane * V);
#pragma unroll
            for (int k = 0; k < V; ++k) bits[k] = w.v[k];
        } else {
#pragma unroll
            for (int k = 0; k < V; ++k) bits[k] = 0u;
        }
        interior_vec<S, C, V, TZ>(out, cur, prev, r, ox + lane * V, oy + row, oz, l2, c2, bits, zlo, zhi);
    }
}

// Two steps of one tile whose cells and one-cell halo are all interior cells
// (none is a boundary, PML or source/receiver cell or lies on the grid edge).
// Step n+1 runs on the tile and its halo, z-plane by z-plane, from planes of
// p^n staged in LDS (a ring of four, each covering the tile plus two cells);
// step n+2 runs on the tile from the planes of p^{n+1} left in LDS (a ring of
// three).  p^{n+1} of the tile goes to out, p^{n+2} to out2.  Arithmetic and
// sum order are interior_column's, and p^{n+1} is rounded to the storage
// type before step n+2 reads it, as a stored value would be, so both steps
// are bit-identical to two single steps.  HBM traffic: p^n and p^{n-1} read
// once, two planes written: 16 B per two FP32 cell updates instead of 24.
template <int TX, int TY> struct PairLds {
    static constexpr int EX = TX + 2, EY = TY + 2, RX = TX + 4, RY = TY + 4;
    static constexpr int NE = EX * EY, NR = RX * RY;
    static constexpr int size = TY >= 2 ? 4 * NR + 3 * NE : 0;  // elements; no pair path for one-row tiles
};

template <typename S, typename C, int TX, int TY, int TZ>
__device__ __forceinline__ void interior_pair(S* __restrict__ o1, S* __restrict__ o2, const S* __restrict__ cur,
                                              const S* __restrict__ prev, const RoomDev& r, int x0, int y0, int z0,
                                              C l2, C c2, C* __restrict__ lds)
{
    using L = PairLds<TX, TY>;
    constexpr int NT = TX * TY;
    constexpr int EX = L::EX, RX = L::RX, NE = L::NE, NR = L::NR;
    constexpr int KE = (NE + NT - 1) / NT, KR = (NR + NT - 1) / NT;
    C* sp = lds;           // [4][NR]: p^n, plane z in slot z & 3
    C* sq = lds + 4 * NR;  // [3][NE]: p^{n+1}, plane z in slot z % 3
    const int64_t sy = r.nx, sz = (int64_t)r.nx * r.ny;
    const int64_t rb = r.base + (int64_t)(y0 - 2) * sy + (x0 - 2);  // plane origin of the p^n rectangle
    const int64_t eb = r.base + (int64_t)(y0 - 1) * sy + (x0 - 1);  // plane origin of tile + halo
    const int tid = (int)threadIdx.x;
    int roff[KR], eoff[KE], erc[KE];
    int64_t rg[KR], eg[KE];
#pragma unroll
    for (int k = 0; k < KR; ++k) {  // rectangle cells of this thread (the last ones clamp and are not stored)
        const int c = tid + k * NT;
        const int cc = c < NR ? c : NR - 1;
        roff[k] = c < NR ? c : -1;
        rg[k] = rb + (int64_t)(cc / RX) * sy + (cc % RX);
    }
#pragma unroll
    for (int k = 0; k < KE; ++k) {  // tile + halo cells of this thread
        const int c = tid + k * NT;
        const int cc = c < NE ? c : NE - 1;
        eoff[k] = c < NE ? c : -1;
        const int ex = cc % EX, ey = cc / EX;
        eg[k] = eb + (int64_t)ey * sy + ex;
        erc[k] = (ey + 1) * RX + (ex + 1);
    }
    const int tx = tid % TX, ty = tid / TX;
    const int ec = (ty + 1) * EX + (tx + 1);  // own cell in a p^{n+1} plane
    const int rc = (ty + 2) * RX + (tx + 2);  // own cell in a p^n plane
    const int64_t og = r.base + (int64_t)(y0 + ty) * sy + (x0 + tx);
    // p^n planes z0-2 and z0-1 to LDS; plane z0 of p^n and plane z0-1 of
    // p^{n-1} into registers (each loop pass prefetches one plane ahead)
    C rp[KR], rm[KE];
#pragma unroll
    for (int k = 0; k < KR; ++k) {
        if (roff[k] >= 0) {
            sp[((z0 - 2) & 3) * NR + roff[k]] = ld<S, C>(cur, (int64_t)(z0 - 2) * sz + rg[k]);
            sp[((z0 - 1) & 3) * NR + roff[k]] = ld<S, C>(cur, (int64_t)(z0 - 1) * sz + rg[k]);
        }
        rp[k] = ld<S, C>(cur, (int64_t)z0 * sz + rg[k]);
    }
#pragma unroll
    for (int k = 0; k < KE; ++k) rm[k] = ld<S, C>(prev, (int64_t)(z0 - 1) * sz + eg[k]);
    for (int z = z0 - 1; z <= z0 + TZ; ++z) {
        // p^n plane z+1 to LDS; prefetch p^n plane z+2 and p^{n-1} plane z+1
        // (clamped to the last planes the tile reads)
        C pm[KE];
#pragma unroll
        for (int k = 0; k < KE; ++k) pm[k] = rm[k];
#pragma unroll
        for (int k = 0; k < KR; ++k)
            if (roff[k] >= 0) sp[((z + 1) & 3) * NR + roff[k]] = rp[k];
        const int zn = z + 2 <= z0 + TZ + 1 ? z + 2 : z0 + TZ + 1;
#pragma unroll
        for (int k = 0; k < KR; ++k) rp[k] = ld<S, C>(cur, (int64_t)zn * sz + rg[k]);
        const int zm = z + 1 <= z0 + TZ ? z + 1 : z0 + TZ;
#pragma unroll
        for (int k = 0; k < KE; ++k) rm[k] = ld<S, C>(prev, (int64_t)zm * sz + eg[k]);
        __syncthreads();
        // step n+1 at plane z over tile + halo
        const C* P0 = sp + (z & 3) * NR;
        const C* Pu = sp + ((z + 1) & 3) * NR;
        const C* Pd = sp + ((z - 1) & 3) * NR;
        C* Q = sq + (z % 3) * NE;
#pragma unroll
        for (int k = 0; k < KE; ++k) {
            if (eoff[k] < 0) continue;
            const int e = erc[k];
            C s = P0[e + 1] + P0[e - 1];
            s = s + P0[e + RX];
            s = s + P0[e - RX];
            s = s + Pu[e];
            s = s + Pd[e];
            const C v = l2 * s + c2 * P0[e];
            Q[eoff[k]] = rt<S, C>(v - pm[k]);
        }
        __syncthreads();
        if (z >= z0 && z < z0 + TZ) st<S, C>(o1, (int64_t)z * sz + og, Q[ec]);
        // step n+2 at plane z-1 over the tile
        if (z - 1 >= z0) {
            const int zc = z - 1;
            const C* Q0 = sq + (zc % 3) * NE;
            const C* Qd = sq + ((zc + 2) % 3) * NE;
            C s = Q0[ec + 1] + Q0[ec - 1];
            s = s + Q0[ec + EX];
            s = s + Q0[ec - EX];
            s = s + Q[ec];
            s = s + Qd[ec];
            const C v = l2 * s + c2 * Q0[ec];
            st<S, C>(o2, (int64_t)zc * sz + og, v - sp[(zc & 3) * NR + rc]);
        }
    }
}

template <typename S, int TX, int TY, int TZ, int V>
__global__ __launch_bounds__(TX / V * TY) void k_interior(S* __restrict__ out, const S* __restrict__ cur,
                                                          const S* __restrict__ prev, const int4* __restrict__ tiles,
                                                          const RoomDev* __restrict__ rooms,
                                                          typename Compute<S>::T l2, typename Compute<S>::T c2)
{
    using C = typename Compute<S>::T;
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;
    const RoomDev r = rooms[t.x];
    tile_interior<S, C, TX, TY, TZ, V>(out, cur, prev, r, t.y, t.z, t.w, l2, c2, nullptr);
}

// ---------------------------------------------------------------- boundary
template <typename C> struct BndScalars {
    C lam_b, l2b, c2b, l2, c2;
};

// Leapfrog value of one cell, same operation order as interior_column.
template <typename S, typename C>
__device__ __forceinline__ C leapfrog_value(const S* __restrict__ cur, const S* __restrict__ prev, int64_t i,
                                            int64_t sy, int64_t sz, C l2, C c2)
{
    C s = ld<S, C>(cur, i + 1) + ld<S, C>(cur, i - 1);
    s = s + ld<S, C>(cur, i + sy);
    s = s + ld<S, C>(cur, i - sy);
    s = s + ld<S, C>(cur, i + sz);
    s = s + ld<S, C>(cur, i - sz);
    const C v = l2 * s + c2 * ld<S, C>(cur, i);
    return v - ld<S, C>(prev, i);
}

// Per-material quotients of the boundary update that depend on the material
// and the boundary scalars only, written by k_mat_consts with the same
// operations bnd_value used to evaluate per entry (so the results are
// bit-identical): {l2b / b0, l * a0 / b0, a0 / (l * b0), 1 / a0}.
constexpr int kMatConsts = 4;

template <typename C>
__global__ void k_mat_consts(const C* __restrict__ A, const C* __restrict__ B, int N, int nmat, C l, C l2,
                             C* __restrict__ M)
{
    const int mat = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (mat >= nmat) return;
    const C a0 = A[(int64_t)mat * N], b0 = B[(int64_t)mat * N];
    C* mk = M + (int64_t)kMatConsts * mat;
    mk[0] = l2 / b0;
    mk[1] = l * a0 / b0;
    mk[2] = a0 / (l * b0);
    mk[3] = (C)1.0 / a0;
}

// New pressure of boundary entry e (rounded to storage precision) with its
// IIR rings advanced.  Mirrored neighbour sum: axes x, y, z in order; a
// solid face replaces its pair by 2*opposite (lrs_1.cl:135-175).  ixb is the
// step's ring index, below N (ring_index).  Memory order: p^n at the cell and
// its six neighbours (boundary cells lie in [1, n-2] on every axis, so all
// six are cells of the room) and p^{n-1} at the cell are requested first,
// then each material's coefficients and ring slots kRingBatch at a time, so
// a material's ring reads are in flight together; the arithmetic and its
// order are those of the reference's update.
constexpr int kRingBatch = 4;

template <typename S, typename C>
__device__ C bnd_value(int e, int64_t i, uint32_t info, uint32_t mats, const S* __restrict__ cur,
                       const S* __restrict__ prev, C* __restrict__ rin, C* __restrict__ rout, int64_t rs,
                       const C* __restrict__ A, const C* __restrict__ B, const C* __restrict__ M, int N,
                       unsigned ixb, int compat, int64_t sy, int64_t sz, const BndScalars<C>& k)
{
    const uint32_t m = info & kInfoMask;
    const int K = (int)((info >> kInfoKShift) & 3u);
    const unsigned sf = (compat && K == 3) ? (unsigned)N / 2u : (unsigned)N;  // lrs_3.cl:70
    const int64_t st3[3] = {1, sy, sz};
    const C pc = ld<S, C>(cur, i);
    const C pd1 = ld<S, C>(prev, i);
    C np[3], nm[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        np[d] = ld<S, C>(cur, i + st3[d]);
        nm[d] = ld<S, C>(cur, i - st3[d]);
    }
    C b0[3], g[3], mq[3][kMatConsts];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        if (q >= K) break;
        const int mat = (int)((mats >> (10 * q)) & 1023u);
        const C* ak = A + mat * N;
        const C* bk = B + mat * N;
        b0[q] = bk[0];
#pragma unroll
        for (int u = 0; u < kMatConsts; ++u) mq[q][u] = M[kMatConsts * mat + u];
        C gn = 0;
        for (int i0 = 1; i0 < N; i0 += kRingBatch) {
            C vb[kRingBatch], vi[kRingBatch], va[kRingBatch], vo[kRingBatch];
#pragma unroll
            for (int w = 0; w < kRingBatch; ++w) {
                const int ii = i0 + w;
                // (ixb + ii) mod sf; ixb + ii < 2N, so one subtraction when sf == N
                unsigned t = ixb + (unsigned)(ii < N ? ii : 1);
                if (sf == (unsigned)N)
                    t = t >= sf ? t - sf : t;
                else
                    t = t % sf;
                const int64_t at = (int64_t)(t + (unsigned)q * sf) * rs + e;
                const int iic = ii < N ? ii : 0;
                vb[w] = bk[iic];
                vi[w] = rin[at];
                va[w] = ak[iic];
                vo[w] = rout[at];
            }
#pragma unroll
            for (int w = 0; w < kRingBatch; ++w)
                if (i0 + w < N) gn += vb[w] * vi[w] - va[w] * vo[w];
        }
        g[q] = gn;
    }
    const C l2 = k.l2b;
    C v;
    if (info & kInfoNoUpdate) {
        // leapfrog_value's operations
        C s = np[0] + nm[0];
        s = s + np[1];
        s = s + nm[1];
        s = s + np[2];
        s = s + nm[2];
        const C lv = k.l2 * s + k.c2 * pc;
        v = lv - pd1;
    } else {
        C acc = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const uint32_t plus = 1u << (2 * d), minus = 2u << (2 * d);
            if (m & plus) {
                const C tt = (C)2.0 * nm[d];
                acc = d == 0 ? tt : acc + tt;
            } else if (m & minus) {
                const C tt = (C)2.0 * np[d];
                acc = d == 0 ? tt : acc + tt;
            } else {
                acc = d == 0 ? np[d] : acc + np[d];
                acc = acc + nm[d];
            }
        }
        C num, den;
        // mq[q][0] = l2b / b0, mq[q][1] = lam_b * a0 / b0
        if (K == 1) {
            num = l2 * acc + k.c2b * pc;
            num = num + mq[0][0] * g[0];
            num = num + (mq[0][1] - (C)1.0) * pd1;
            den = (C)1.0 + mq[0][1];
        } else {
            const int U = (compat && K == 3) ? 2 : K;  // lrs_3.cl:96-174 use materials 0 and 1
            C G = g[0] / b0[0];
            C Asum = mq[0][1];
            den = (C)1.0 + mq[0][1];
            for (int q = 1; q < U; ++q) {
                G = G + g[q] / b0[q];
                Asum = Asum + mq[q][1];
                den = den + mq[q][1];
            }
            num = l2 * acc + l2 * G;
            num = num + k.c2b * pc;
            num = num + (Asum - (C)1.0) * pd1;
        }
        v = num / den;
    }
    v = rt<S, C>(v);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        if (q >= K) break;
        // mq[q][2] = a0 / (lam_b * b0), mq[q][3] = 1 / a0
        const C xn = mq[q][2] * (v - pd1) - g[q] / b0[q];
        const unsigned w = ixb + (unsigned)q * sf;
        rin[w * rs + e] = xn;
        rout[w * rs + e] = mq[q][3] * (b0[q] * xn + g[q]);
    }
    return v;
}

__device__ __forceinline__ unsigned ring_index(int64_t t, int N)
{
    // simulation.hpp:365 decrements the ring index before every step's LRS
    // launches, starting from 0: step t (0-based) uses (-(t+1)) mod N.
    const int64_t r = (t + 1) % N;
    return (unsigned)((N - r) % N);
}

template <typename S, typename C>
struct BndCtx {
    S* out;
    const S* cur;
    const S* prev;
    S* out2;          // p^{n+2} of two-step tiles (work kind 4)
    const S* keep;    // the buffer whose value edge-plane special cells keep
    const RoomDev* rooms;
    RoomDev room0;
    const int32_t* e_cell;
    const uint32_t* e_info;
    const uint32_t* e_mat;
    C* rin;
    C* rout;
    int64_t rs;
    const C* A;
    const C* B;
    const C* M;       // per-material quotients (k_mat_consts)
    int N;
    int compat;
    BndScalars<C> k;
    // special
    const int32_t* sp_cell;
    const int32_t* sp_room;
    const int32_t* sp_bidx;
    const int32_t* sp_src_off;
    const int32_t* sp_src;
    const int32_t* sp_mic_off;
    const int32_t* sp_mic;
    const float* src;
    int64_t src_cap;
    const int64_t* src_len;
    float* traces;
    int64_t trace_cap;
    C scale, inv_scale;
    PmlDev<C> pml;           // split path (single phi/psi buffers)
    // fused path PML: phi/psi double buffers selected by step parity,
    // zeta at cell centres and +1/2 points, and the PML cell list
    C* phib[2][3];
    C* psib[2];
    const C* zh[3];
    C cc, idt, idx;
    // PML cell records {x | y << 11 | z << 22, word, k, nb}: word bits 24-31
    // room, bits 0-23 kind + 2 (kind -2: boundary cell, phi/psi only; -1:
    // plain; >= 0: special-cell index); k the cell's compact index; nb bits
    // 0-10 / 11-21 the x / y extent of its box, bits 22-27 which of the
    // neighbours +x, -x, +y, -y, +z, -z lie in the same box, bit 28 set when
    // one of the others lies in another PML box: then pml_side[q] holds the
    // box (bits 0-25) and those neighbours (bits 26-31), found by pml_find;
    // bit 29 set when at least two of the cell's zeta are non-zero (only
    // then does the fused PML chunk read and store the cell's psi)
    const int4* pml_rec;
    const uint32_t* pml_side;
    const int32_t* pml_boxes;
    const int4* pml_blk;
    const uint32_t* pml_bmask;
};

__device__ __forceinline__ void unpack_xyz(uint32_t p, int& x, int& y, int& z)
{
    x = (int)(p & 2047u);
    y = (int)((p >> 11) & 2047u);
    z = (int)(p >> 22);
}

template <typename S, typename C>
__device__ __forceinline__ void boundary_entry(const BndCtx<S, C>& c, int e, unsigned ixb, bool skip_special)
{
    const uint32_t info = c.e_info[e];
    if (skip_special && (info & kInfoSpecial)) return;
    const uint32_t room = info >> kInfoRoomShift;
    const RoomDev r = room == 0 ? c.room0 : c.rooms[room];
    const int64_t i = c.e_cell[e];
    const int64_t sy = r.nx, sz = (int64_t)r.nx * r.ny;
    const C v = bnd_value<S, C>(e, i, info, c.e_mat[e], c.cur, c.prev, c.rin, c.rout, c.rs, c.A, c.B, c.M, c.N, ixb,
                                c.compat, sy, sz, c.k);
    st<S, C>(c.out, i, v);
}

// Boundary entry from its packed work-order record {cell, info, mats, entry}.
template <typename S, typename C>
__device__ __forceinline__ void boundary_packed(const BndCtx<S, C>& c, int4 en, unsigned ixb)
{
    const uint32_t info = (uint32_t)en.y;
    if (info & kInfoSpecial) return;
    const uint32_t room = info >> kInfoRoomShift;
    const RoomDev r = room == 0 ? c.room0 : c.rooms[room];
    const int64_t i = en.x;
    const int64_t sy = r.nx, sz = (int64_t)r.nx * r.ny;
    const C v = bnd_value<S, C>(en.w, i, info, (uint32_t)en.z, c.cur, c.prev, c.rin, c.rout, c.rs, c.A, c.B, c.M, c.N, ixb,
                                c.compat, sy, sz, c.k);
    st<S, C>(c.out, i, v);
}

// Adds special cell j's sources (in order) to its new value v, stores it and
// records its receivers (source.cl:32-40 then mic.cl:32-41).
template <typename S, typename C>
__device__ __forceinline__ void finish_special(const BndCtx<S, C>& c, int j, int64_t i, int64_t t, C v)
{
    for (int q = c.sp_src_off[j]; q < c.sp_src_off[j + 1]; ++q) {
        const int s = c.sp_src[q];
        const C sv = t < c.src_len[s] ? (C)c.src[(int64_t)s * c.src_cap + t] : (C)0;
        v = v + sv * c.inv_scale;
    }
    st<S, C>(c.out, i, v);
    const C stored = ld<S, C>(c.out, i);
    for (int q = c.sp_mic_off[j]; q < c.sp_mic_off[j + 1]; ++q) {
        const int m = c.sp_mic[q];
        c.traces[(int64_t)m * c.trace_cap + t] = (float)(stored * c.scale);
    }
}

// mode 1: full update (boundary or leapfrog value) + sources + receivers
// mode 2: add sources to the stored value + receivers (after a PML pass)
// mode 3: as mode 1 inside the fused launch; PML cells (b == -3) are left to
//         the PML threads, which finish them
template <typename S, typename C>
__device__ __forceinline__ void special_entry(const BndCtx<S, C>& c, int j, unsigned ixb, int64_t t, int mode)
{
    const int64_t i = c.sp_cell[j];
    const int b = c.sp_bidx[j];
    const bool full = mode == 1 || mode == 3;
    if (mode == 3 && b == -3) return;
    C v;
    if (full && b >= 0) {
        const uint32_t info = c.e_info[b];
        const uint32_t room = info >> kInfoRoomShift;
        const RoomDev r = room == 0 ? c.room0 : c.rooms[room];
        v = bnd_value<S, C>(b, i, info, c.e_mat[b], c.cur, c.prev, c.rin, c.rout, c.rs, c.A, c.B, c.M, c.N, ixb,
                            c.compat, (int64_t)r.nx, (int64_t)r.nx * r.ny, c.k);
    } else if (full && b == -1) {
        const RoomDev r = c.rooms[c.sp_room[j]];
        v = rt<S, C>(leapfrog_value<S, C>(c.cur, c.prev, i, (int64_t)r.nx, (int64_t)r.nx * r.ny, c.k.l2, c.k.c2));
    } else {
        // edge planes (b == -2) and add-only mode keep the stored value: the
        // output buffer's, or with two steps per pass, for the second step,
        // the buffer a single step would have written into
        v = ld<S, C>(c.keep, i);
    }
    finish_special<S, C>(c, j, i, t, v);
}

// ---------------------------------------------------------------- fused PML
// Inside the fused launch of step n the auxiliary field phi^n is produced
// from phi^{n-1}, p^n (cur), p^{n-1} (prev) and psi^{n-1/2}: the update the
// split path runs after step n-1, moved into the next launch.  Each PML cell
// stores its own phi^n and recomputes phi^n at its -x, -y, -z neighbours for
// the divergence, so no cell waits on another.
// phi^n of component d at cell (x, y, z) = global index j of room r, from
// phi^{n-1}, p^n, p^{n-1} and psi^{n-1/2}: kj is the cell's compact index
// (-1: outside the PML region, where phi is zero), kn that of its +
// neighbour along d (-1: psi is zero there).
template <typename S, typename C>
__device__ __forceinline__ C phi_at(const BndCtx<S, C>& c, const RoomDev& r, int d, int64_t j, int x, int y, int z,
                                    int64_t kj, int64_t kn, const C* __restrict__ phr, const C* __restrict__ psr,
                                    int64_t st)
{
    if (kj < 0) return (C)0;
    const C* zc0 = c.pml.z[0] + r.axis_off[0];
    const C* zc1 = c.pml.z[1] + r.axis_off[1];
    const C* zc2 = c.pml.z[2] + r.axis_off[2];
    C h, g2, g3;
    if (d == 0) {
        h = c.zh[0][r.axis_off[0] + x];
        g2 = zc1[y] + zc2[z] - h;
        g3 = zc1[y] * zc2[z];
    } else if (d == 1) {
        h = c.zh[1][r.axis_off[1] + y];
        g2 = zc0[x] + zc2[z] - h;
        g3 = zc0[x] * zc2[z];
    } else {
        h = c.zh[2][r.axis_off[2] + z];
        g2 = zc0[x] + zc1[y] - h;
        g3 = zc0[x] * zc1[y];
    }
    return phi_update<C>(phr[kj], ld<S, C>(c.cur, j), ld<S, C>(c.cur, j + st), ld<S, C>(c.prev, j),
                         ld<S, C>(c.prev, j + st), psr[kj], kn >= 0 ? psr[kn] : (C)0, h, g2, g3, c.cc, c.idt,
                         c.idx);
}

// One PML cell of the fused launch in two parts.  pml_first: phi^n at the
// cell (stored), psi^{n+1/2} (stored).  pml_second: unless it is a boundary
// cell, the PML pressure from phi^n at the -x, -y, -z neighbours, given
// (have[d]) or recomputed.
template <typename C> struct PmlCell {
    RoomDev r;
    int q, flag, x, y, z;
    uint32_t nb;
    int64_t i, k;
    C fx, fy, fz, pc, ps, psn;
};

// Compact index of the neighbour of PML cell u in direction dir (0..5: +x,
// -x, +y, -y, +z, -z), or -1 outside the PML region.
template <typename S, typename C>
__device__ __forceinline__ int64_t pml_nb(const BndCtx<S, C>& c, const PmlCell<C>& u, int dir)
{
    const int64_t wx = u.nb & 2047u, wxy = wx * ((u.nb >> 11) & 2047u);
    const int64_t step = (dir >> 1) == 0 ? 1 : ((dir >> 1) == 1 ? wx : wxy);
    if ((u.nb >> (22 + dir)) & 1u) return (dir & 1) ? u.k - step : u.k + step;
    if (!((u.nb >> 28) & 1u)) return -1;
    const uint32_t side = c.pml_side[u.q];
    if (!((side >> (26 + dir)) & 1u)) return -1;
    const int sgn = (dir & 1) ? -1 : 1;
    return pml_find(c.pml_boxes, u.r, (int)(side & 0x3FFFFFFu), u.x + ((dir >> 1) == 0 ? sgn : 0),
                    u.y + ((dir >> 1) == 1 ? sgn : 0), u.z + ((dir >> 1) == 2 ? sgn : 0));
}

template <typename S, typename C, bool kStore = true>
__device__ __forceinline__ void pml_first(const BndCtx<S, C>& c, int q, int par, PmlCell<C>& u)
{
    const int4 rec = c.pml_rec[q];
    const uint32_t word = (uint32_t)rec.y;
    u.q = q;
    u.r = c.rooms[word >> 24];
    u.flag = (int)(word & 0xFFFFFFu) - 2;
    unpack_xyz((uint32_t)rec.x, u.x, u.y, u.z);
    u.k = (uint32_t)rec.z;
    u.nb = (uint32_t)rec.w;
    const int64_t sy = u.r.nx, sz = (int64_t)u.r.nx * u.r.ny;
    u.i = u.r.base + (int64_t)u.z * sz + (int64_t)u.y * sy + u.x;
    const C* psr = c.psib[par ^ 1];
    const int64_t kxp = pml_nb<S, C>(c, u, 0);
    const int64_t kyp = pml_nb<S, C>(c, u, 2);
    const int64_t kzp = pml_nb<S, C>(c, u, 4);
    u.fx = phi_at<S, C>(c, u.r, 0, u.i, u.x, u.y, u.z, u.k, kxp, c.phib[par ^ 1][0], psr, 1);
    u.fy = phi_at<S, C>(c, u.r, 1, u.i, u.x, u.y, u.z, u.k, kyp, c.phib[par ^ 1][1], psr, sy);
    u.fz = phi_at<S, C>(c, u.r, 2, u.i, u.x, u.y, u.z, u.k, kzp, c.phib[par ^ 1][2], psr, sz);
    if (kStore) {
        c.phib[par][0][u.k] = u.fx;
        c.phib[par][1][u.k] = u.fy;
        c.phib[par][2][u.k] = u.fz;
    }
    u.pc = ld<S, C>(c.cur, u.i);
    u.ps = psr[u.k];
    u.psn = u.ps + c.pml.dt * u.pc;
    if (kStore) c.psib[par][u.k] = u.psn;
}

template <typename S, typename C>
__device__ __forceinline__ void pml_second(const BndCtx<S, C>& c, const PmlCell<C>& u, int par, int64_t t, C fxm,
                                           C fym, C fzm, bool hx, bool hy, bool hz)
{
    const int64_t sy = u.r.nx, sz = (int64_t)u.r.nx * u.r.ny;
    const C* psr = c.psib[par ^ 1];
    if (!hx)
        fxm = phi_at<S, C>(c, u.r, 0, u.i - 1, u.x - 1, u.y, u.z, pml_nb<S, C>(c, u, 1), u.k, c.phib[par ^ 1][0], psr, 1);
    if (!hy)
        fym = phi_at<S, C>(c, u.r, 1, u.i - sy, u.x, u.y - 1, u.z, pml_nb<S, C>(c, u, 3), u.k, c.phib[par ^ 1][1], psr,
                           sy);
    if (!hz)
        fzm = phi_at<S, C>(c, u.r, 2, u.i - sz, u.x, u.y, u.z - 1, pml_nb<S, C>(c, u, 5), u.k, c.phib[par ^ 1][2], psr,
                           sz);
    const C dv = (u.fx - fxm) + (u.fy - fym) + (u.fz - fzm);
    const int64_t i = u.i;
    C sum = ld<S, C>(c.cur, i + 1) + ld<S, C>(c.cur, i - 1);
    sum = sum + ld<S, C>(c.cur, i + sy);
    sum = sum + ld<S, C>(c.cur, i - sy);
    sum = sum + ld<S, C>(c.cur, i + sz);
    sum = sum + ld<S, C>(c.cur, i - sz);
    const C v = pml_formula<C>(sum, u.pc, ld<S, C>(c.prev, i), dv, u.ps, u.psn, c.pml.z[0][u.r.axis_off[0] + u.x],
                               c.pml.z[1][u.r.axis_off[1] + u.y], c.pml.z[2][u.r.axis_off[2] + u.z], c.pml.dt,
                               c.pml.dx, c.k.l2);
    if (u.flag >= 0)
        finish_special<S, C>(c, u.flag, i, t, rt<S, C>(v));
    else
        st<S, C>(c.out, i, v);
}

// PML cell q on its own: phi^n at the -x, -y, -z neighbours is recomputed.
template <typename S, typename C>
__device__ __forceinline__ void pml_entry(const BndCtx<S, C>& c, int q, int64_t t)
{
    const int par = (int)(t & 1);
    PmlCell<C> u;
    pml_first<S, C>(c, q, par, u);
    if (u.flag == -2) return;  // boundary cell: its pressure comes from the LRS update
    pml_second<S, C>(c, u, par, t, (C)0, (C)0, (C)0, false, false, false);
}

// PML cells [a, b) of one work chunk, one per thread of the workgroup.  Each
// thread shares its phi^n through LDS; a neighbour's phi^n is read from there
// when the neighbour is in the chunk (nbd[q]: list-index distances, 10 bits
// per axis, 0 = not in the chunk) and recomputed otherwise.  The arithmetic
// is pml_first's and pml_second's; the order of the memory accesses is not:
// every value the cell needs (its own phi^n inputs, those of the -x, -y, -z
// neighbours, the 7-point sum of p^n) is requested in one round, with no
// load behind a branch or a store, before anything is computed or stored.
// Indexes that point outside the PML region (compact index -1) are replaced
// by the cell's own for the load and the value is then discarded.
template <typename S, typename C, int NT, bool kStore = true>
__device__ __forceinline__ void pml_chunk(const BndCtx<S, C>& c, int a, int b, int64_t t,
                                          const uint32_t* __restrict__ nbd, C (*lds)[NT])
{
    const int tid = (int)threadIdx.x;
    const int q = a + tid;
    const bool valid = q < b;
    const int par = (int)(t & 1);
    const C* psr = c.psib[par ^ 1];
    int flag = -2;
    int64_t i = 0;
    C fx = 0, fy = 0, fz = 0, rxm = 0, rym = 0, rzm = 0;
    C sum = 0, pc = 0, pm = 0, ps = 0, psn = 0, sx = 0, syy = 0, szz = 0;
    uint32_t nbq = 0;
    if (valid) {
        PmlCell<C> u;
        const int4 rec = c.pml_rec[q];
        nbq = nbd[q];
        const uint32_t word = (uint32_t)rec.y;
        const uint32_t room = word >> 24;
        u.q = q;
        u.r = room == 0 ? c.room0 : c.rooms[room];
        flag = (int)(word & 0xFFFFFFu) - 2;
        unpack_xyz((uint32_t)rec.x, u.x, u.y, u.z);
        u.k = (uint32_t)rec.z;
        u.nb = (uint32_t)rec.w;
        const RoomDev& r = u.r;
        const int64_t sy = r.nx, sz = (int64_t)r.nx * r.ny;
        const int64_t st[3] = {1, sy, sz};
        i = r.base + (int64_t)u.z * sz + (int64_t)u.y * sy + u.x;
        const int64_t k = u.k;
        int64_t kp[3], km[3];  // compact indexes of the + and - neighbours
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            kp[d] = pml_nb<S, C>(c, u, 2 * d);
            km[d] = pml_nb<S, C>(c, u, 2 * d + 1);
        }
        // ---- loads
        const C* zc[3] = {c.pml.z[0] + r.axis_off[0], c.pml.z[1] + r.axis_off[1], c.pml.z[2] + r.axis_off[2]};
        const int co[3] = {u.x, u.y, u.z};
        C zeta[3], hp[3], hm[3], oldp[3], oldm[3], np1[3], nm1[3], npv[3], nmv[3], psp[3], psm[3];
        pc = ld<S, C>(c.cur, i);
        pm = ld<S, C>(c.prev, i);
        // psi enters the update only through products with zeta_a * zeta_b
        // (a != b) taken at this cell's coordinates, so where fewer than two
        // of the cell's zeta are non-zero (record bit 29 clear: a cell of a
        // one-face slab) those products are zero whatever psi holds, and psi
        // is neither read nor stored.
        const bool live = (u.nb >> 29) & 1u;
        ps = 0;
        if (live) ps = psr[k];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const C* ph = c.phib[par ^ 1][d];
            const int64_t kpd = kp[d] >= 0 ? kp[d] : k;
            const int64_t kmd = km[d] >= 0 ? km[d] : k;
            zeta[d] = zc[d][co[d]];
            hp[d] = c.zh[d][r.axis_off[d] + co[d]];
            hm[d] = c.zh[d][r.axis_off[d] + co[d] - 1];
            oldp[d] = ph[k];
            oldm[d] = ph[kmd];
            np1[d] = ld<S, C>(c.cur, i + st[d]);
            nm1[d] = ld<S, C>(c.cur, i - st[d]);
            npv[d] = ld<S, C>(c.prev, i + st[d]);
            nmv[d] = ld<S, C>(c.prev, i - st[d]);
            psp[d] = 0;
            psm[d] = 0;
            if (live) {
                psp[d] = psr[kpd];
                psm[d] = psr[kmd];
            }
        }
        // ---- phi^n at the cell (phi_at with kj = k, kn = kp[d]) and at its
        // - neighbours (kj = km[d], kn = k); g2, g3 as phi_at forms them
        C f[3], g[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int d1 = d == 0 ? 1 : 0, d2 = d == 2 ? 1 : 2;
            const C g2p = zeta[d1] + zeta[d2] - hp[d];
            const C g2m = zeta[d1] + zeta[d2] - hm[d];
            const C g3 = zeta[d1] * zeta[d2];
            f[d] = phi_update<C>(oldp[d], pc, np1[d], pm, npv[d], ps, kp[d] >= 0 ? psp[d] : (C)0, hp[d], g2p, g3,
                                 c.cc, c.idt, c.idx);
            g[d] = km[d] < 0 ? (C)0
                             : phi_update<C>(oldm[d], nm1[d], pc, nmv[d], pm, psm[d], ps, hm[d], g2m, g3, c.cc,
                                             c.idt, c.idx);
        }
        fx = f[0];
        fy = f[1];
        fz = f[2];
        rxm = g[0];
        rym = g[1];
        rzm = g[2];
        psn = ps + c.pml.dt * pc;
        sum = np1[0] + nm1[0];
        sum = sum + np1[1];
        sum = sum + nm1[1];
        sum = sum + np1[2];
        sum = sum + nm1[2];
        sx = zeta[0];
        syy = zeta[1];
        szz = zeta[2];
        // ---- stores
        if (kStore) {
            c.phib[par][0][k] = fx;
            c.phib[par][1][k] = fy;
            c.phib[par][2][k] = fz;
            if (live) c.psib[par][k] = psn;
        }
        lds[0][tid] = fx;
        lds[1][tid] = fy;
        lds[2][tid] = fz;
    }
    __syncthreads();
    if (!valid || flag == -2) return;
    const int dxm = (int)(nbq & 1023u), dym = (int)((nbq >> 10) & 1023u), dzm = (int)(nbq >> 20);
    const C fxm = dxm ? lds[0][tid - dxm] : rxm;
    const C fym = dym ? lds[1][tid - dym] : rym;
    const C fzm = dzm ? lds[2][tid - dzm] : rzm;
    const C dv = (fx - fxm) + (fy - fym) + (fz - fzm);
    const C v = pml_formula<C>(sum, pc, pm, dv, ps, psn, sx, syy, szz, c.pml.dt, c.pml.dx, c.k.l2);
    if (flag >= 0)
        finish_special<S, C>(c, flag, i, t, rt<S, C>(v));
    else
        st<S, C>(c.out, i, v);
}

// One PML block of the fused launch: an ex x ey block of columns of one PML
// box, each thread one column marching ez planes of z.  Per plane every
// thread computes phi^n (three components) and psi^{n+1/2} at its cell and
// stores them, exactly as pml_first does from a record; after one barrier it
// takes phi^n of its -x and -y neighbours from the threads that computed them
// (LDS, two planes double-buffered), of its -z neighbour from its own
// previous plane, and recomputes only those that lie outside the block, as
// pml_second does.  Cells whose bit dz is set in the column's skip mask
// (boundary cells, whose pressure is the LRS update, and source/receiver
// cells, finished from their record) get phi and psi but no pressure here.
template <typename S, typename C, int NT>
__device__ __forceinline__ void pml_block(const BndCtx<S, C>& c, int blk, int64_t t, C* __restrict__ lds)
{
    const int4 bd = c.pml_blk[blk];
    const int32_t* B = c.pml_boxes + 8 * bd.x;
    const RoomDev r = c.rooms[B[0]];
    int ox, oy, oz;
    unpack_xyz((uint32_t)bd.y, ox, oy, oz);
    const int ex = bd.z & 255, ey = (bd.z >> 8) & 1023, ez = (int)((uint32_t)bd.z >> 18);
    const int tid = (int)threadIdx.x;
    const int tx = tid % ex, ty = tid / ex;
    const bool active = ty < ey;
    const int x = ox + tx, y = oy + (active ? ty : 0);
    const int wx = B[2] - B[1], wy = B[4] - B[3];
    const int64_t wxy = (int64_t)wx * wy;
    const int64_t sy = r.nx, sz = (int64_t)r.nx * r.ny;
    int64_t k = (int64_t)B[7] + ((int64_t)(oz - B[5]) * wy + (y - B[3])) * wx + (x - B[1]);
    int64_t i = r.base + (int64_t)oz * sz + (int64_t)y * sy + x;
    const uint32_t skip = active ? c.pml_bmask[bd.w + tid] : ~0u;
    const int par = (int)(t & 1);
    const C* psr = c.psib[par ^ 1];
    const C* phx = c.phib[par ^ 1][0];
    const C* phy = c.phib[par ^ 1][1];
    const C* phz = c.phib[par ^ 1][2];
    C fzc = (C)0;  // phi_z^n one plane down in this column
    for (int dz = 0; dz < ez; ++dz) {
        const int z = oz + dz;
        C* L = lds + (dz & 1) * 2 * NT;
        C fx = (C)0, fy = (C)0, fz = (C)0, pc = (C)0, ps = (C)0, psn = (C)0;
        if (active) {
            const int64_t kxp = x + 1 < B[2] ? k + 1 : pml_find(c.pml_boxes, r, bd.x, x + 1, y, z);
            const int64_t kyp = y + 1 < B[4] ? k + wx : pml_find(c.pml_boxes, r, bd.x, x, y + 1, z);
            const int64_t kzp = z + 1 < B[6] ? k + wxy : pml_find(c.pml_boxes, r, bd.x, x, y, z + 1);
            fx = phi_at<S, C>(c, r, 0, i, x, y, z, k, kxp, phx, psr, 1);
            fy = phi_at<S, C>(c, r, 1, i, x, y, z, k, kyp, phy, psr, sy);
            fz = phi_at<S, C>(c, r, 2, i, x, y, z, k, kzp, phz, psr, sz);
            c.phib[par][0][k] = fx;
            c.phib[par][1][k] = fy;
            c.phib[par][2][k] = fz;
            pc = ld<S, C>(c.cur, i);
            ps = psr[k];
            psn = ps + c.pml.dt * pc;
            c.psib[par][k] = psn;
            L[tid] = fx;
            L[NT + tid] = fy;
        }
        __syncthreads();
        if (active && !((skip >> dz) & 1u)) {
            C fxm, fym, fzm;
            if (tx > 0) {
                fxm = L[tid - 1];
            } else {
                const int64_t kxm = x - 1 >= B[1] ? k - 1 : pml_find(c.pml_boxes, r, bd.x, x - 1, y, z);
                fxm = phi_at<S, C>(c, r, 0, i - 1, x - 1, y, z, kxm, k, phx, psr, 1);
            }
            if (ty > 0) {
                fym = L[NT + tid - ex];
            } else {
                const int64_t kym = y - 1 >= B[3] ? k - wx : pml_find(c.pml_boxes, r, bd.x, x, y - 1, z);
                fym = phi_at<S, C>(c, r, 1, i - sy, x, y - 1, z, kym, k, phy, psr, sy);
            }
            if (dz > 0) {
                fzm = fzc;
            } else {
                const int64_t kzm = z - 1 >= B[5] ? k - wxy : pml_find(c.pml_boxes, r, bd.x, x, y, z - 1);
                fzm = phi_at<S, C>(c, r, 2, i - sz, x, y, z - 1, kzm, k, phz, psr, sz);
            }
            const C dv = (fx - fxm) + (fy - fym) + (fz - fzm);
            C sum = ld<S, C>(c.cur, i + 1) + ld<S, C>(c.cur, i - 1);
            sum = sum + ld<S, C>(c.cur, i + sy);
            sum = sum + ld<S, C>(c.cur, i - sy);
            sum = sum + ld<S, C>(c.cur, i + sz);
            sum = sum + ld<S, C>(c.cur, i - sz);
            const C v = pml_formula<C>(sum, pc, ld<S, C>(c.prev, i), dv, ps, psn, c.pml.z[0][r.axis_off[0] + x],
                                       c.pml.z[1][r.axis_off[1] + y], c.pml.z[2][r.axis_off[2] + z], c.pml.dt,
                                       c.pml.dx, c.k.l2);
            st<S, C>(c.out, i, v);
        }
        fzc = fz;
        k += wxy;
        i += sz;
    }
}

template <typename S, typename C>
__host__ BndCtx<S, C> make_ctx(const StepArgs& a)
{
    BndCtx<S, C> c;
    c.out = (S*)a.out;
    c.cur = (const S*)a.cur;
    c.prev = (const S*)a.prev;
    c.out2 = (S*)a.out2;
    c.keep = a.keep ? (const S*)a.keep : (const S*)a.out;
    c.rooms = a.rooms;
    c.room0 = a.room0;
    c.e_cell = a.e_cell;
    c.e_info = a.e_info;
    c.e_mat = a.e_mat;
    c.rin = (C*)a.ring_in;
    c.rout = (C*)a.ring_out;
    c.rs = a.nb;
    c.A = (const C*)a.coef_a;
    c.B = (const C*)a.coef_b;
    c.M = (const C*)a.coef_m;
    c.N = a.ncoef;
    c.compat = a.compat;
    c.k.lam_b = (C)a.lam_b;
    c.k.l2b = (C)a.l2b;
    c.k.c2b = (C)a.c2b;
    c.k.l2 = (C)a.l2;
    c.k.c2 = (C)a.c2;
    c.sp_cell = a.sp_cell;
    c.sp_room = a.sp_room;
    c.sp_bidx = a.sp_bidx;
    c.sp_src_off = a.sp_src_off;
    c.sp_src = a.sp_src;
    c.sp_mic_off = a.sp_mic_off;
    c.sp_mic = a.sp_mic;
    c.src = a.src;
    c.src_cap = a.src_cap;
    c.src_len = a.src_len;
    c.traces = a.traces;
    c.trace_cap = a.trace_cap;
    c.scale = (C)a.scale;
    c.inv_scale = (C)(1.0 / a.scale);
    c.pml.phx = (C*)a.pml_phi[0][0];
    c.pml.phy = (C*)a.pml_phi[0][1];
    c.pml.phz = (C*)a.pml_phi[0][2];
    c.pml.psi = (C*)a.pml_psi[0];
    for (int d = 0; d < 3; ++d) {
        c.pml.z[d] = (const C*)a.pml_z[d];
        c.zh[d] = (const C*)a.pml_zh[d];
        c.phib[0][d] = (C*)a.pml_phi[0][d];
        c.phib[1][d] = (C*)a.pml_phi[1][d];
    }
    c.psib[0] = (C*)a.pml_psi[0];
    c.psib[1] = (C*)a.pml_psi[1];
    c.pml.dt = (C)a.dt;
    c.pml.dx = (C)a.dx;
    const C cv = (C)a.c;
    c.cc = cv * cv;
    c.idt = (C)1.0 / c.pml.dt;
    c.idx = (C)1.0 / c.pml.dx;
    c.pml_rec = a.pml_rec;
    c.pml_side = a.pml_side;
    c.pml_boxes = a.pml_boxes;
    c.pml_blk = a.pml_blk;
    c.pml_bmask = a.pml_bmask;
    return c;
}

constexpr int kBndThreads = 256;

template <typename S, typename C>
__global__ __launch_bounds__(kBndThreads) void k_boundary(BndCtx<S, C> c, int nb, int nblk_b, int nsp,
                                                          const int64_t* __restrict__ d_t, int t_rel,
                                                          int skip_special, int special_mode)
{
    const int64_t t = *d_t + t_rel;
    const unsigned ixb = c.N > 0 ? ring_index(t, c.N) : 0u;
    if ((int)blockIdx.x < nblk_b) {
        const int e = blockIdx.x * kBndThreads + threadIdx.x;
        if (e < nb) boundary_entry<S, C>(c, e, ixb, skip_special != 0);
    } else {
        const int j = (blockIdx.x - nblk_b) * kBndThreads + threadIdx.x;
        if (j < nsp && special_mode != 0) special_entry<S, C>(c, j, ixb, t, special_mode);
    }
}

// ---------------------------------------------------------------- fused step
// One launch per step.  Workgroups of four kinds never touch the same cell
// (interior tiles skip the cells flagged in `mask`):
//   [0, nblk_b)            boundary entries, one per thread
//   [.., +nblk_s)          source/receiver cells
//   [.., +nblk_p)          PML cells (phi, psi and pressure)
//   [tile0, tile0+ntiles)  interior tiles; tile0 is a multiple of 8 so the
//                          XCD-swizzled tile order keeps its XCD mapping
// The latency-bound lists come first so they overlap the interior stream.
// Minimum waves per SIMD the fused kernel is compiled for.  With 16-bit
// storage and several cells per lane the register budget of seven waves (at
// most 72 VGPRs) fits every work kind without spilling, and the 640x480x320
// FP16 step runs 3.7 % faster (280.8 against 291.5 us).  Elsewhere the
// compiler's own choice (74 VGPRs for FP32, 102 for FP64) is kept: the same
// cap spills with one cell per lane and is slower for FP32.
template <typename S, int V> struct FusedMinWaves {
    static constexpr int value = (sizeof(S) == 2 && V > 1) ? 7 : 1;
};

template <typename S, int TX, int TY, int TZ, int V>
__global__ __launch_bounds__(TX / V * TY) __attribute__((amdgpu_waves_per_eu(FusedMinWaves<S, V>::value)))
void k_fused(const int4* __restrict__ work, const int64_t* __restrict__ d_t, int t_rel,
             const int4* __restrict__ e_pack, const uint32_t* __restrict__ pml_nbd,
             const uint32_t* __restrict__ colmask, BndCtx<S, typename Compute<S>::T> c)
{
    // The small arguments come first: the build preloads the first kernel
    // argument dwords into SGPRs, so the work item's address is known at wave
    // start instead of after a load from the argument segment.
    using C = typename Compute<S>::T;
    constexpr int NT = TX / V * TY;   // threads
    constexpr int NCOL = TX * TY;     // columns (column-mask words per tile slot)
    constexpr int kPair = V == 1 ? PairLds<TX, TY>::size : 0;
    constexpr int kLds = kPair > 4 * NT ? kPair : 4 * NT;
    __shared__ C lds[kLds];  // PML chunks: phi^n by thread; two-step tiles: pressure planes
    // Work table: item (kind, a, b, w).  kind 0: interior tile of slot a,
    // origin packed in b (x | y << 11 | z << 22), room w; 4: the same tile
    // advanced two steps (interior_pair); 1: packed boundary entries
    // e_pack[a..b); 2: PML list [a..b); 3: special cells [a..b); 5: PML
    // block a (pml_block); 6: PML list [a..b) of source/receiver cells
    // whose phi and psi a PML block stores; -1: idle.
    // Items of one XCD (blockIdx % 8) follow its tiles in order, each list
    // chunk right before the tile holding its first cell, so both read the
    // same planes of p^n through L2.
    // the step index is read here, in the same round as the work item, not
    // after it: list workgroups would otherwise wait on one more dependent load
    const int64_t t = *d_t + t_rel;
    const int4 wk = work[blockIdx.x];
    if constexpr (kPair == 0) {
        if (wk.x == 4) return;  // no two-step tiles are dealt to these shapes
    } else {
        if (wk.x == 4) {
            const RoomDev r = wk.w == 0 ? c.room0 : c.rooms[wk.w];
            const uint32_t o = (uint32_t)wk.z;
            interior_pair<S, C, TX, TY, TZ>(c.out, c.out2, c.cur, c.prev, r, (int)(o & 2047u),
                                            (int)((o >> 11) & 2047u), (int)(o >> 22), c.k.l2, c.k.c2, lds);
            return;
        }
    }
    if (wk.x == 0) {
        // w: room (bits 0-15), then the tile's first and end plane of this
        // item (bits 16-23, 24-31; an end of 0 is the whole tile)
        const int room = wk.w & 0xFFFF, zlo = (wk.w >> 16) & 0xFF, zend = (int)((uint32_t)wk.w >> 24);
        const RoomDev r = room == 0 ? c.room0 : c.rooms[room];
        const uint32_t o = (uint32_t)wk.z;
        tile_interior<S, C, TX, TY, TZ, V>(c.out, c.cur, c.prev, r, (int)(o & 2047u), (int)((o >> 11) & 2047u),
                                           (int)(o >> 22), c.k.l2, c.k.c2, colmask + (int64_t)wk.y * NCOL, zlo,
                                           zend ? zend : TZ);
        return;
    }
    if (wk.x < 0) return;
    if (wk.x == 2) {  // whole workgroup: it meets a barrier
        pml_chunk<S, C, NT>(c, wk.y, wk.z, t, pml_nbd, reinterpret_cast<C(*)[NT]>(lds));
        return;
    }
    if (wk.x == 5) {  // whole workgroup: barriers every plane
        pml_block<S, C, NT>(c, wk.y, t, lds);
        return;
    }
    if (wk.x == 6) {  // PML source/receiver cells next to PML blocks, which store their phi and psi
        pml_chunk<S, C, NT, false>(c, wk.y, wk.z, t, pml_nbd, reinterpret_cast<C(*)[NT]>(lds));
        return;
    }
    const int q = wk.y + (int)threadIdx.x;
    if (q >= wk.z) return;
    const unsigned ixb = c.N > 0 ? ring_index(t, c.N) : 0u;
    if (wk.x == 1)
        boundary_packed<S, C>(c, e_pack[q], ixb);
    else
        special_entry<S, C>(c, q, ixb, t, 3);
}

// The same step with the lists in leading workgroups of their own
// (AFX_LIST_BLOCKS):
//   [0, nblk_b)            boundary entries, one per thread
//   [.., +nblk_s)          source/receiver cells
//   [.., +nblk_p)          PML cells (phi, psi and pressure)
//   [tile0, tile0+ntiles)  interior tiles; tile0 is a multiple of 8 so the
//                          XCD-swizzled tile order keeps its XCD mapping
template <typename S, int TX, int TY, int TZ, int V>
__global__ __launch_bounds__(TX / V * TY) void k_fused_blocks(BndCtx<S, typename Compute<S>::T> c,
                                                          const int4* __restrict__ tiles,
                                                          const uint32_t* __restrict__ colmask, int nb, int nblk_b,
                                                          int nsp, int nblk_s, int npml, int nblk_p, int tile0,
                                                          const int64_t* __restrict__ d_t, int t_rel)
{
    using C = typename Compute<S>::T;
    constexpr int NT = TX / V * TY;   // threads
    constexpr int NCOL = TX * TY;     // columns
    const int bid = (int)blockIdx.x;
    if (bid >= tile0) {
        const int slot = bid - tile0;
        const int4 t4 = tiles[slot];
        if (t4.x < 0) return;
        const RoomDev r = c.rooms[t4.x];
        tile_interior<S, C, TX, TY, TZ, V>(c.out, c.cur, c.prev, r, t4.y, t4.z, t4.w, c.k.l2, c.k.c2,
                                           colmask + (int64_t)slot * NCOL);
        return;
    }
    const int64_t t = *d_t + t_rel;
    if (bid < nblk_b) {
        const unsigned ixb = c.N > 0 ? ring_index(t, c.N) : 0u;
        const int e = bid * NT + (int)threadIdx.x;
        if (e < nb) boundary_entry<S, C>(c, e, ixb, true);
    } else if (bid < nblk_b + nblk_s) {
        const unsigned ixb = c.N > 0 ? ring_index(t, c.N) : 0u;
        const int j = (bid - nblk_b) * NT + (int)threadIdx.x;
        if (j < nsp) special_entry<S, C>(c, j, ixb, t, 3);
    } else if (bid < nblk_b + nblk_s + nblk_p) {
        const int k = (bid - nblk_b - nblk_s) * NT + (int)threadIdx.x;
        if (k < npml) pml_entry<S, C>(c, k, t);
    }
}

// Plain copy, the calibration kernel of afx_copy_probe.
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ a, float4* __restrict__ b, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

hipError_t launch_copy(const void* src, void* dst, int64_t n16, hipStream_t s)
{
    AFX_CHECK_LAUNCH(dim3(8192), dim3(256), 256);
    hipLaunchKernelGGL(k_copy, dim3(8192), dim3(256), 0, s, (const float4*)src, (float4*)dst, n16);
    return hipGetLastError();
}

__global__ void k_advance_t(int64_t* d_t, int64_t k)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *d_t += k;
}

// ---------------------------------------------------------------- PML, physical
// Grote-Sim PML on the slab boxes.  psi is cell-centred at half steps; phi_x
// lives at x+1/2 and is stored at index x (likewise y, z).  Pressure part
// runs after the interior launch, before the boundary/special launch.
template <typename S>
__global__ void k_pml_p(PmlArgs a)
{
    using C = typename Compute<S>::T;
    // grid: x*y flattened on blockIdx.x, z plane on blockIdx.y, box on blockIdx.z
    const int32_t* bx = a.boxes + 8 * blockIdx.z;
    const int room = bx[0];
    const int wx = bx[2] - bx[1], wy = bx[4] - bx[3];
    const RoomDev r = a.rooms[room];
    const int64_t sy = r.nx, sz = (int64_t)r.nx * r.ny;
    S* out = (S*)a.out;
    const S* cur = (const S*)a.cur;
    const S* prev = (const S*)a.prev;
    PmlDev<C> pq;
    pq.phx = (C*)a.phi[0];
    pq.phy = (C*)a.phi[1];
    pq.phz = (C*)a.phi[2];
    pq.psi = (C*)a.psi;
    pq.dt = (C)a.dt;
    pq.dx = (C)a.dx;
    const C* zx = (const C*)a.z[0] + r.axis_off[0];
    const C* zy = (const C*)a.z[1] + r.axis_off[1];
    const C* zz = (const C*)a.z[2] + r.axis_off[2];
    const C l2 = (C)a.l2;
    {
        const int rr = (int)(blockIdx.x * blockDim.x + threadIdx.x);
        const int z = bx[5] + (int)blockIdx.y;
        if (rr >= wx * wy || z >= bx[6]) return;
        const int x = bx[1] + rr % wx;
        const int y = bx[3] + rr / wx;
        const int64_t i = r.base + (int64_t)z * sz + (int64_t)y * sy + x;
        const int64_t k = (int64_t)bx[7] + (int64_t)(z - bx[5]) * wx * wy + rr;
        const int hint = (int)blockIdx.z;
        const int64_t kxm = pml_find(a.boxes, r, hint, x - 1, y, z);
        const int64_t kym = pml_find(a.boxes, r, hint, x, y - 1, z);
        const int64_t kzm = pml_find(a.boxes, r, hint, x, y, z - 1);
        const C pc = ld<S, C>(cur, i);
        const C pm = ld<S, C>(prev, i);
        C sum = ld<S, C>(cur, i + 1) + ld<S, C>(cur, i - 1);
        sum = sum + ld<S, C>(cur, i + sy);
        sum = sum + ld<S, C>(cur, i - sy);
        sum = sum + ld<S, C>(cur, i + sz);
        sum = sum + ld<S, C>(cur, i - sz);
        st<S, C>(out, i, pml_value<C>(sum, pc, pm, k, kxm, kym, kzm, zx[x], zy[y], zz[z], pq, l2));
    }
}

// Auxiliary phi update; runs after the step's pressure is final.
template <typename S>
__global__ void k_pml_phi(PmlArgs a)
{
    using C = typename Compute<S>::T;
    // grid: x*y flattened on blockIdx.x, z plane on blockIdx.y, box on blockIdx.z
    const int32_t* bx = a.boxes + 8 * blockIdx.z;
    const int room = bx[0];
    const int wx = bx[2] - bx[1], wy = bx[4] - bx[3];
    const RoomDev r = a.rooms[room];
    const int64_t sy = r.nx, sz = (int64_t)r.nx * r.ny;
    const S* p1 = (const S*)a.out;
    const S* p = (const S*)a.cur;
    C* phx = (C*)a.phi[0];
    C* phy = (C*)a.phi[1];
    C* phz = (C*)a.phi[2];
    const C* psi = (const C*)a.psi;
    const C* zx = (const C*)a.z[0] + r.axis_off[0];
    const C* zy = (const C*)a.z[1] + r.axis_off[1];
    const C* zz = (const C*)a.z[2] + r.axis_off[2];
    const C* hx = (const C*)a.zh[0] + r.axis_off[0];
    const C* hy = (const C*)a.zh[1] + r.axis_off[1];
    const C* hz = (const C*)a.zh[2] + r.axis_off[2];
    const C c = (C)a.c, dx = (C)a.dx, dt = (C)a.dt;
    const C cc = c * c, idt = (C)1.0 / dt, idx = (C)1.0 / dx;
    {
        const int rr = (int)(blockIdx.x * blockDim.x + threadIdx.x);
        const int z = bx[5] + (int)blockIdx.y;
        if (rr >= wx * wy || z >= bx[6]) return;
        const int x = bx[1] + rr % wx;
        const int y = bx[3] + rr / wx;
        const int64_t i = r.base + (int64_t)z * sz + (int64_t)y * sy + x;
        const int64_t k = (int64_t)bx[7] + (int64_t)(z - bx[5]) * wx * wy + rr;
        const int hint = (int)blockIdx.z;
        const int64_t kxp = pml_find(a.boxes, r, hint, x + 1, y, z);
        const int64_t kyp = pml_find(a.boxes, r, hint, x, y + 1, z);
        const int64_t kzp = pml_find(a.boxes, r, hint, x, y, z + 1);
        const C p1c = ld<S, C>(p1, i), pcc = ld<S, C>(p, i);
        {
            const C h = hx[x];
            phx[k] = phi_update<C>(phx[k], p1c, ld<S, C>(p1, i + 1), pcc, ld<S, C>(p, i + 1), psi[k],
                                   kxp >= 0 ? psi[kxp] : (C)0, h, zy[y] + zz[z] - h, zy[y] * zz[z], cc, idt, idx);
        }
        {
            const C h = hy[y];
            phy[k] = phi_update<C>(phy[k], p1c, ld<S, C>(p1, i + sy), pcc, ld<S, C>(p, i + sy), psi[k],
                                   kyp >= 0 ? psi[kyp] : (C)0, h, zx[x] + zz[z] - h, zx[x] * zz[z], cc, idt, idx);
        }
        {
            const C h = hz[z];
            phz[k] = phi_update<C>(phz[k], p1c, ld<S, C>(p1, i + sz), pcc, ld<S, C>(p, i + sz), psi[k],
                                   kzp >= 0 ? psi[kzp] : (C)0, h, zx[x] + zy[y] - h, zx[x] * zy[y], cc, idt, idx);
        }
    }
}

// ---------------------------------------------------------------- PML, compat
// pml.cl:154-265 over 1 < x < n-2 (likewise y, z).  The reference reads p and
// psiph at +1 neighbours while other work-items write them; here the three
// parts run as ordered phases (psi, p, fi), each reading the completed
// previous phase.  Only room 0 (compat is single-room).
template <typename S, int PHASE>
__global__ void k_pml_compat(PmlArgs a)
{
    using C = typename Compute<S>::T;
    const RoomDev r = a.rooms[0];
    const int nx = r.nx, ny = r.ny, nz = r.nz;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    const int64_t wx = nx - 4, wy = ny - 4, wz = nz - 4;
    if (wx <= 0 || wy <= 0 || wz <= 0) return;
    const int64_t total = wx * wy * wz;
    S* p = (S*)a.out;
    const S* pm1 = (const S*)a.cur;
    const S* pm2 = (const S*)a.prev;
    C* fx = (C*)a.phi[0];
    C* fy = (C*)a.phi[1];
    C* fz = (C*)a.phi[2];
    const C* fxm1 = (const C*)a.phi_old[0];
    const C* fym1 = (const C*)a.phi_old[1];
    const C* fzm1 = (const C*)a.phi_old[2];
    C* psiph = (C*)a.psi;
    const C* psimh = (const C*)a.psi_old;
    const C* kx = (const C*)a.z[0];
    const C* ky = (const C*)a.z[1];
    const C* kz = (const C*)a.z[2];
    const C c = (C)a.c, dx = (C)a.dx, dt = (C)a.dt;
    const int L = a.width;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        const int x = 2 + (int)(q % wx);
        const int y = 2 + (int)((q / wx) % wy);
        const int z = 2 + (int)(q / (wx * wy));
        const int64_t i = r.base + (int64_t)z * sz + (int64_t)y * sy + x;
        if (PHASE == 0) {
            psiph[i] = dt * ld<S, C>(pm1, i) + psimh[i];
        } else if (PHASE == 1) {
            auto P1 = [&](int64_t j) { return ld<S, C>(pm1, j); };
            C lap = P1(i + 1) + P1(i - 1) + P1(i + sy) + P1(i - sy) + P1(i + sz) + P1(i - sz) - (C)6.0 * P1(i);
            auto av = [&](const C* f, int64_t i0, int64_t da, int64_t db) {
                return (C)0.25 * (f[i0] + f[i0 + db] + f[i0 + da] + f[i0 + da + db]);
            };
            C dv = av(fxm1, i + 1, sy, sz) - av(fxm1, i, sy, sz) + av(fym1, i + sy, 1, sz) - av(fym1, i, 1, sz)
                   + av(fzm1, i + sz, 1, sy) - av(fzm1, i, 1, sy);
            const C t1 = c * c / (dx * dx) * lap;
            const C t2 = (C)1.0 / dx * dv;
            const C t3 = kx[x] * ky[y] * kz[z] * (C)0.5 * (psiph[i] + psimh[i]);
            const C t4 = (C)1.0 / (dt * dt) * (ld<S, C>(pm2, i) - (C)2.0 * P1(i));
            const C t5 = (kx[x] * ky[y] + ky[y] * kz[z] + kz[z] * kx[x]) * P1(i);
            const C t6 = (kx[x] + ky[y] + kz[z]) / ((C)2.0 * dt) * ld<S, C>(pm2, i);
            const C num = t1 + t2 - t3 - t4 - t5 + t6;
            st<S, C>(p, i, num / ((C)1.0 / (dt * dt) + (kx[x] + ky[y] + kz[z]) / ((C)2.0 * dt)));
        } else {
            auto avs = [&](const S* f, int64_t i0, int64_t da, int64_t db) {
                return (C)0.25 * (ld<S, C>(f, i0) + ld<S, C>(f, i0 + db) + ld<S, C>(f, i0 + da) + ld<S, C>(f, i0 + da + db));
            };
            auto avc = [&](const C* f, int64_t i0, int64_t da, int64_t db) {
                return (C)0.25 * (f[i0] + f[i0 + db] + f[i0 + da] + f[i0 + da + db]);
            };
            const C kax = (C)0.5 * (kx[x] + kx[x + 1]);
            const C kay = (C)0.5 * (ky[y] + ky[y + 1]);
            const C kaz = (C)0.5 * (kz[z] + kz[z + 1]);
            if (x < L || x > nx - L - 1) {
                const C d = (C)0.5 / dx * (avs(p, i + 1, sy, sz) - avs(p, i, sy, sz) + avs(pm1, i + 1, sy, sz) - avs(p, i, sy, sz));
                const C dp = (C)1.0 / dx * (avc(psiph, i + 1, sy, sz) - avc(psiph, i, sy, sz));
                fx[i] = (-(C)0.5 * kax * fxm1[i] + (kay + kaz - kax) * d + (kay * kaz) * dp + (C)1.0 / dt * (fxm1[i]))
                        / ((C)1.0 / dt + kax * (C)0.5);
            }
            if (y < L || y > ny - L - 1) {
                const C d = (C)0.5 / dx * (avs(p, i + sy, 1, sz) - avs(p, i, 1, sz) + avs(pm1, i + sy, 1, sz) - avs(p, i, 1, sz));
                const C dp = (C)1.0 / dx * (avc(psiph, i + sy, 1, sz) - avc(psiph, i, 1, sz));
                const C kxy = (C)0.5 * (kx[y] + kx[y + 1]);  // ksi_avg(ksi_x, y), pml.cl:247
                fy[i] = (-(C)0.5 * kay * fym1[i] + (kax + kaz - kay) * d + (kxy * kaz) * dp + (C)1.0 / dt * (fym1[i]))
                        / ((C)1.0 / dt + kay * (C)0.5);
            }
            if (z < L || z > nz - L - 1) {
                const C d = (C)0.5 / dx * (avs(p, i + sz, 1, sy) - avs(p, i, 1, sy) + avs(pm1, i + sz, 1, sy) - avs(p, i, 1, sy));
                const C dp = (C)1.0 / dx * (avc(psiph, i + sz, 1, sy) - avc(psiph, i, 1, sy));
                fz[i] = (-(C)0.5 * kaz * fzm1[i] + (kax + kay - kaz) * d + (kax * kay) * dp + (C)1.0 / dt * (fzm1[i]))
                        / ((C)1.0 / dt + kaz * (C)0.5);
            }
        }
    }
}

// ---------------------------------------------------------------- utilities
__global__ void k_fill_zero(uint32_t* p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

template <typename S>
__global__ void k_to_f32(const S* __restrict__ src, float* __restrict__ dst, int64_t n, double scale)
{
    using C = typename Compute<S>::T;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (float)((C)src[i] * (C)scale);
}

// One axis-aligned plane of a field, as FP32: dst[v][u] = src[base + u*su + v*sv] * scale.
// copy_plane.cl:29-34 does the y-plane (u = x, v = z) for the live view.
template <typename S>
__global__ void k_plane_f32(const S* __restrict__ src, float* __restrict__ dst, int nu, int nv, int64_t su,
                            int64_t sv, int64_t base, double scale)
{
    using C = typename Compute<S>::T;
    const int64_t n = (int64_t)nu * nv;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t u = k % nu, v = k / nu;
        dst[k] = (float)((C)src[base + u * su + v * sv] * (C)scale);
    }
}

hipError_t launch_plane_f32(const void* src, Prec prec, float* dst, int nu, int nv, int64_t su, int64_t sv,
                            int64_t base, double scale, hipStream_t s)
{
    int64_t blocks = ((int64_t)nu * nv + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) return hipSuccess;
    const dim3 g((unsigned)blocks), b(256);
    AFX_CHECK_LAUNCH(g, b, 1024);
    switch (prec) {
    case Prec::F32: hipLaunchKernelGGL(k_plane_f32<float>, g, b, 0, s, (const float*)src, dst, nu, nv, su, sv, base, scale); break;
    case Prec::F64: hipLaunchKernelGGL(k_plane_f32<double>, g, b, 0, s, (const double*)src, dst, nu, nv, su, sv, base, scale); break;
    case Prec::F16: hipLaunchKernelGGL(k_plane_f32<_Float16>, g, b, 0, s, (const _Float16*)src, dst, nu, nv, su, sv, base, scale); break;
    case Prec::BF16: hipLaunchKernelGGL(k_plane_f32<__bf16>, g, b, 0, s, (const __bf16*)src, dst, nu, nv, su, sv, base, scale); break;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------- launchers
template <typename S, int TX, int TY, int TZ, int V = 1>
static hipError_t interior_t(const StepArgs& a, hipStream_t s)
{
    using C = typename Compute<S>::T;
    constexpr int NT = TX / V * TY;
    AFX_CHECK_LAUNCH(dim3(a.ntiles), dim3(NT), NT);
    hipLaunchKernelGGL((k_interior<S, TX, TY, TZ, V>), dim3(a.ntiles), dim3(NT), 0, s, (S*)a.out,
                       (const S*)a.cur, (const S*)a.prev, a.tiles, a.rooms, (C)a.l2, (C)a.c2);
    return hipGetLastError();
}

template <typename S, int TX, int TY, int TZ, int V = 1>
static hipError_t fused_t(const StepArgs& a, hipStream_t s)
{
    using C = typename Compute<S>::T;
    constexpr int NT = TX / V * TY;
    BndCtx<S, C> c = make_ctx<S, C>(a);
    if (a.work) {
        AFX_CHECK_LAUNCH(dim3(a.nwork), dim3(NT), NT);
        hipLaunchKernelGGL((k_fused<S, TX, TY, TZ, V>), dim3(a.nwork), dim3(NT), 0, s, a.work, a.d_t, a.t_rel,
                           a.e_pack, a.pml_nbd, a.colmask, c);
        return hipGetLastError();
    }
    const int nblk_b = (a.nb + NT - 1) / NT;
    const int nblk_s = (a.nsp + NT - 1) / NT;
    const int nblk_p = (a.npml + NT - 1) / NT;
    const int tile0 = (nblk_b + nblk_s + nblk_p + 7) / 8 * 8;
    AFX_CHECK_LAUNCH(dim3(tile0 + a.ntiles), dim3(NT), NT);
    hipLaunchKernelGGL((k_fused_blocks<S, TX, TY, TZ, V>), dim3(tile0 + a.ntiles), dim3(NT), 0, s, c, a.tiles,
                       a.colmask, a.nb, nblk_b, a.nsp, nblk_s, a.npml, nblk_p, tile0, a.d_t, a.t_rel);
    return hipGetLastError();
}

template <typename S>
static hipError_t interior_s(const StepArgs& a, hipStream_t s)
{
    switch (a.tile) {
    case 0: return interior_t<S, 64, 4, 16>(a, s);
    case 1: return interior_t<S, 64, 4, 32>(a, s);
    case 2: return interior_t<S, 64, 8, 16>(a, s);
    case 3: return interior_t<S, 64, 2, 32>(a, s);
    case 4: return interior_t<S, 128, 2, 16>(a, s);
    case 5: return interior_t<S, 64, 4, 8>(a, s);
    case 6: return interior_t<S, 256, 1, 16>(a, s);
    case 7: return interior_t<S, 256, 4, 8, 4>(a, s);
    case 8: return interior_t<S, 256, 2, 8, 4>(a, s);
    case 9: return interior_t<S, 128, 4, 16, 2>(a, s);
    case 10: return interior_t<S, 128, 4, 8, 2>(a, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename S>
static hipError_t fused_s(const StepArgs& a, hipStream_t s)
{
    switch (a.tile) {
    case 0: return fused_t<S, 64, 4, 16>(a, s);
    case 1: return fused_t<S, 64, 4, 32>(a, s);
    case 2: return fused_t<S, 64, 8, 16>(a, s);
    case 3: return fused_t<S, 64, 2, 32>(a, s);
    case 4: return fused_t<S, 128, 2, 16>(a, s);
    case 5: return fused_t<S, 64, 4, 8>(a, s);
    case 6: return fused_t<S, 256, 1, 16>(a, s);
    case 7: return fused_t<S, 256, 4, 8, 4>(a, s);
    case 8: return fused_t<S, 256, 2, 8, 4>(a, s);
    case 9: return fused_t<S, 128, 4, 16, 2>(a, s);
    case 10: return fused_t<S, 128, 4, 8, 2>(a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_interior(const StepArgs& a, hipStream_t s)
{
    if (a.ntiles <= 0) return hipSuccess;
    switch (a.prec) {
    case Prec::F32: return interior_s<float>(a, s);
    case Prec::F64: return interior_s<double>(a, s);
    case Prec::F16: return interior_s<_Float16>(a, s);
    case Prec::BF16: return interior_s<__bf16>(a, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_fused(const StepArgs& a, hipStream_t s)
{
    if (a.ntiles <= 0) return hipSuccess;
    switch (a.prec) {
    case Prec::F32: return fused_s<float>(a, s);
    case Prec::F64: return fused_s<double>(a, s);
    case Prec::F16: return fused_s<_Float16>(a, s);
    case Prec::BF16: return fused_s<__bf16>(a, s);
    }
    return hipErrorInvalidValue;
}

template <typename S>
static hipError_t boundary_s(const StepArgs& a, int skip_special, int special_mode, hipStream_t s)
{
    using C = typename Compute<S>::T;
    const int nblk_b = (a.nb + kBndThreads - 1) / kBndThreads;
    const int nsp = special_mode ? a.nsp : 0;
    const int nblk_s = (nsp + kBndThreads - 1) / kBndThreads;
    if (nblk_b + nblk_s == 0) return hipSuccess;
    BndCtx<S, C> c = make_ctx<S, C>(a);
    AFX_CHECK_LAUNCH(dim3(nblk_b + nblk_s), dim3(kBndThreads), kBndThreads);
    hipLaunchKernelGGL((k_boundary<S, C>), dim3(nblk_b + nblk_s), dim3(kBndThreads), 0, s, c, a.nb, nblk_b, nsp,
                       a.d_t, a.t_rel, skip_special, special_mode);
    return hipGetLastError();
}

hipError_t launch_boundary(const StepArgs& a, int skip_special, int special_mode, hipStream_t s)
{
    switch (a.prec) {
    case Prec::F32: return boundary_s<float>(a, skip_special, special_mode, s);
    case Prec::F64: return boundary_s<double>(a, skip_special, special_mode, s);
    case Prec::F16: return boundary_s<_Float16>(a, skip_special, special_mode, s);
    case Prec::BF16: return boundary_s<__bf16>(a, skip_special, special_mode, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_advance_t(int64_t* d_t, int64_t k, hipStream_t s)
{
    AFX_CHECK_LAUNCH(dim3(1), dim3(64), 1024);
    hipLaunchKernelGGL(k_advance_t, dim3(1), dim3(64), 0, s, d_t, k);
    return hipGetLastError();
}

static dim3 pml_grid(const PmlArgs& a)
{
    const int blocks = (a.box_xy_max + 255) / 256;
    return dim3((unsigned)(blocks > 0 ? blocks : 1), (unsigned)(a.box_z_max > 0 ? a.box_z_max : 1), (unsigned)a.nbox);
}

hipError_t launch_pml_phys_p(const PmlArgs& a, hipStream_t s)
{
    if (a.nbox <= 0) return hipSuccess;
    AFX_CHECK_LAUNCH(pml_grid(a), dim3(256), 1024);
    switch (a.prec) {
    case Prec::F32: hipLaunchKernelGGL(k_pml_p<float>, pml_grid(a), dim3(256), 0, s, a); break;
    case Prec::F64: hipLaunchKernelGGL(k_pml_p<double>, pml_grid(a), dim3(256), 0, s, a); break;
    case Prec::F16: hipLaunchKernelGGL(k_pml_p<_Float16>, pml_grid(a), dim3(256), 0, s, a); break;
    case Prec::BF16: hipLaunchKernelGGL(k_pml_p<__bf16>, pml_grid(a), dim3(256), 0, s, a); break;
    }
    return hipGetLastError();
}

hipError_t launch_pml_phys_phi(const PmlArgs& a, hipStream_t s)
{
    if (a.nbox <= 0) return hipSuccess;
    AFX_CHECK_LAUNCH(pml_grid(a), dim3(256), 1024);
    switch (a.prec) {
    case Prec::F32: hipLaunchKernelGGL(k_pml_phi<float>, pml_grid(a), dim3(256), 0, s, a); break;
    case Prec::F64: hipLaunchKernelGGL(k_pml_phi<double>, pml_grid(a), dim3(256), 0, s, a); break;
    case Prec::F16: hipLaunchKernelGGL(k_pml_phi<_Float16>, pml_grid(a), dim3(256), 0, s, a); break;
    case Prec::BF16: hipLaunchKernelGGL(k_pml_phi<__bf16>, pml_grid(a), dim3(256), 0, s, a); break;
    }
    return hipGetLastError();
}

template <typename S>
static hipError_t pml_compat_s(const PmlArgs& a, hipStream_t s)
{
    const dim3 g(2048), b(256);
    AFX_CHECK_LAUNCH(g, b, 1024);
    hipLaunchKernelGGL((k_pml_compat<S, 0>), g, b, 0, s, a);
    hipLaunchKernelGGL((k_pml_compat<S, 1>), g, b, 0, s, a);
    hipLaunchKernelGGL((k_pml_compat<S, 2>), g, b, 0, s, a);
    return hipGetLastError();
}

hipError_t launch_pml_compat(const PmlArgs& a, hipStream_t s)
{
    switch (a.prec) {
    case Prec::F32: return pml_compat_s<float>(a, s);
    case Prec::F64: return pml_compat_s<double>(a, s);
    case Prec::F16: return pml_compat_s<_Float16>(a, s);
    case Prec::BF16: return pml_compat_s<__bf16>(a, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_mat_consts(Prec prec, const void* A, const void* B, int N, int nmat, double lam_b, double l2b,
                             void* M, hipStream_t s)
{
    if (nmat <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)((nmat + 255) / 256);
    AFX_CHECK_LAUNCH(dim3(blocks), dim3(256), 1024);
    if (prec == Prec::F64)
        hipLaunchKernelGGL(k_mat_consts<double>, dim3(blocks), dim3(256), 0, s, (const double*)A, (const double*)B, N,
                           nmat, lam_b, l2b, (double*)M);
    else
        hipLaunchKernelGGL(k_mat_consts<float>, dim3(blocks), dim3(256), 0, s, (const float*)A, (const float*)B, N,
                           nmat, (float)lam_b, (float)l2b, (float*)M);
    return hipGetLastError();
}

hipError_t launch_fill(void* p, int64_t bytes, hipStream_t s)
{
    const int64_t n = bytes / 4;
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    AFX_CHECK_LAUNCH(dim3((unsigned)blocks), dim3(256), 1024);
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(256), 0, s, (uint32_t*)p, n);
    return hipGetLastError();
}

hipError_t launch_to_f32(const void* src, Prec prec, float* dst, int64_t n, double scale, hipStream_t s)
{
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) return hipSuccess;
    AFX_CHECK_LAUNCH(dim3((unsigned)blocks), dim3(256), 1024);
    switch (prec) {
    case Prec::F32: hipLaunchKernelGGL(k_to_f32<float>, dim3((unsigned)blocks), dim3(256), 0, s, (const float*)src, dst, n, scale); break;
    case Prec::F64: hipLaunchKernelGGL(k_to_f32<double>, dim3((unsigned)blocks), dim3(256), 0, s, (const double*)src, dst, n, scale); break;
    case Prec::F16: hipLaunchKernelGGL(k_to_f32<_Float16>, dim3((unsigned)blocks), dim3(256), 0, s, (const _Float16*)src, dst, n, scale); break;
    case Prec::BF16: hipLaunchKernelGGL(k_to_f32<__bf16>, dim3((unsigned)blocks), dim3(256), 0, s, (const __bf16*)src, dst, n, scale); break;
    }
    return hipGetLastError();
}

}  // namespace afx
