// Internal types shared by the host orchestration (afx_host.cpp) and the
// gfx950 kernels (afx_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace afx {

// One room inside the shared field allocation.  base is the offset of cell
// (0,0,0); cells are x-fastest.
struct RoomDev {
    int64_t base;
    int32_t nx, ny, nz;
    int32_t pml;         // 1 when the room has a physical PML
    // a cell is in the PML slabs when x < pml_lo[0] or x >= pml_hi[0] (likewise y, z)
    int32_t pml_lo[3], pml_hi[3];
    int32_t axis_off[3]; // offset of this room's per-axis zeta profiles
    int32_t box0, box1;  // this room's boxes in the physical PML box table
};

// Boundary entry info word:
//   bits 0-5   solid-face mask, bit order X_P, X_M, Y_P, Y_M, Z_P, Z_M
//   bits 8-12  reference group id 0..25 (lrs-data.hpp:53-78)
//   bits 16-17 number of materials K (1..3)
//   bit  20    compat: pressure not updated (group X_M_Y_M_Z_M, simulation.hpp:452)
//   bit  21    the cell carries a source or receiver (handled by the special pass)
//   bits 24-31 room id
constexpr uint32_t kInfoMask = 63u;
constexpr int kInfoGroupShift = 8;
constexpr int kInfoKShift = 16;
constexpr uint32_t kInfoNoUpdate = 1u << 20;
constexpr uint32_t kInfoSpecial = 1u << 21;
constexpr int kInfoRoomShift = 24;
constexpr int kMaxRooms = 256;
constexpr int kMaxMaterials = 1024;  // 10-bit material ids, three per entry word
constexpr int kMaxCoef = 8;

// Interior tile shapes the step kernels are instantiated for: tx x ty x tz
// cells.  vec > 1: each lane owns vec consecutive cells of x (one 64-lane
// wave per row, tx = 64 vec), read and written as one vector; such tiles
// need every room's nx to be a multiple of vec.
struct TileShape {
    int tx, ty, tz, vec;
};
constexpr int kNumTiles = 11;
extern const TileShape kTiles[kNumTiles];

enum class Prec : int { F32 = 0, F64 = 1, F16 = 2, BF16 = 3 };

// Everything a step launch needs.  Pointers are device pointers; the field
// pointers are void* and reinterpreted according to prec.
struct StepArgs {
    Prec prec;
    int tile;                // index into kTiles
    void* out;               // p^{n+1}
    const void* cur;         // p^n
    const void* prev;        // p^{n-1}
    void* out2;              // p^{n+2} of two-step tiles (work kind 4), else null
    const void* keep;        // buffer edge-plane special cells keep (null: out)
    const int4* tiles;       // {room, x0, y0, z0}; room < 0 marks padding
    int32_t ntiles;
    const RoomDev* rooms;
    RoomDev room0;           // rooms[0], by value: tiles of room 0 need no load
    // scalars (double so one struct serves FP32 and FP64 kernels)
    double l2, c2;           // interior lambda^2 and 2(1-3 lambda^2)
    double lam_b, l2b, c2b;  // boundary 'l', l*l, 2(1-3 l*l)
    int compat;
    // boundary list
    int32_t nb;
    const int32_t* e_cell;   // global linear cell index
    const uint32_t* e_info;
    const uint32_t* e_mat;   // three 10-bit global material ids
    void* ring_in;           // [3*ncoef][nb] of compute type
    void* ring_out;
    const void* coef_a;      // [nmat][ncoef] of compute type
    const void* coef_b;
    const void* coef_m;      // [nmat][4] of compute type: l2b/b0, lam_b*a0/b0, a0/(lam_b*b0), 1/a0
    int32_t ncoef;
    // special cells (sources / receivers)
    int32_t nsp;
    const int32_t* sp_cell;
    const int32_t* sp_room;
    const int32_t* sp_bidx;  // boundary entry index or -1
    const int32_t* sp_src_off;  // [nsp+1] into sp_src
    const int32_t* sp_src;      // source ids
    const int32_t* sp_mic_off;  // [nsp+1] into sp_mic
    const int32_t* sp_mic;      // receiver ids
    const float* src;           // [nsrc][src_cap]
    int64_t src_cap;
    const int64_t* src_len;     // per source
    float* traces;              // [nmic][trace_cap]
    int64_t trace_cap;
    double scale;               // stored value = pressure / scale
    // time
    const int64_t* d_t;         // device step counter (start of the chunk)
    int32_t t_rel;              // step offset inside the chunk
    // fused step kernel: per tile slot and tile thread, bit z - z0 marks a
    // cell the tile leaves to a list (boundary, special or PML cell)
    const uint32_t* colmask;
    // physical PML state (compute type): phi and psi double buffers (the
    // split path uses set 0 in place), zeta at cell centres and +1/2 points,
    // and the fused path's PML cell list
    void* pml_phi[2][3];
    void* pml_psi[2];
    const void* pml_z[3];
    const void* pml_zh[3];
    const int4* pml_rec;        // {packed xyz, room << 24 | kind + 2, compact index, box extent/neighbour bits}
    const uint32_t* pml_side;   // box | neighbours in other boxes << 26, for cells flagged in pml_rec
    const int32_t* pml_boxes;   // [nbox][8]: room, x0, x1, y0, y1, z0, z1, compact base
    int32_t npml;
    // fused path, PML blocks: {box, packed origin, ex | ey << 8 | ez << 18,
    // mask offset} per block and one skip mask per block column (bit dz:
    // the pressure of that cell is left to another thread)
    const int4* pml_blk;
    const uint32_t* pml_bmask;
    // fused path work table (kind, a, b) per workgroup, XCD-interleaved:
    // interior tiles with the list chunks holding their cells (boundary
    // entries as packed records {cell, info, mats, entry} grouped by tile;
    // the PML list, itself grouped by tile, with the in-chunk distances of
    // each cell's -x/-y/-z neighbours; special cells); null runs the lists
    // in leading workgroups
    const int4* work;
    int32_t nwork;
    const int4* e_pack;
    const uint32_t* pml_nbd;
    double dt, dx, c;
};

// PML arguments
struct PmlArgs {
    Prec prec;
    void* out; const void* cur; const void* prev;
    void* phi[3];               // compute type: compact PML region (physical) / full-grid new fi (compat)
    void* phi_old[3];           // compat: previous fi
    void* psi;                  // physical psi / compat psiph
    void* psi_old;              // compat psimh
    const void* z[3];           // zeta at cell centres per axis (compute type), global per room offset
    const void* zh[3];          // zeta at +1/2 points
    const int32_t* boxes;       // [nbox][8]: room, x0, x1, y0, y1, z0, z1, compact base of phi/psi
    int32_t nbox;
    int32_t box_xy_max;         // largest box x*y extent (grid sizing)
    int32_t box_z_max;          // largest box z extent
    const RoomDev* rooms;
    int32_t nrooms;
    double l2, c, dx, dt;
    int32_t width;
};

// launchers (afx_kernels.hip)
hipError_t launch_interior(const StepArgs& a, hipStream_t s);
hipError_t launch_boundary(const StepArgs& a, int skip_special, int special_mode, hipStream_t s);
hipError_t launch_fused(const StepArgs& a, hipStream_t s);
hipError_t launch_advance_t(int64_t* d_t, int64_t k, hipStream_t s);
hipError_t launch_pml_phys_p(const PmlArgs& a, hipStream_t s);
hipError_t launch_pml_phys_phi(const PmlArgs& a, hipStream_t s);
hipError_t launch_pml_compat(const PmlArgs& a, hipStream_t s);
hipError_t launch_fill(void* p, int64_t bytes, hipStream_t s);
hipError_t launch_mat_consts(Prec prec, const void* A, const void* B, int N, int nmat, double lam_b, double l2b,
                             void* M, hipStream_t s);
hipError_t launch_copy(const void* src, void* dst, int64_t n16, hipStream_t s);
hipError_t launch_to_f32(const void* src, Prec prec, float* dst, int64_t n, double scale, hipStream_t s);
hipError_t launch_plane_f32(const void* src, Prec prec, float* dst, int nu, int nv, int64_t su, int64_t sv,
                            int64_t base, double scale, hipStream_t s);

}  // namespace afx
