/*
 * afx — MI355X-native FDTD room-acoustics core: the C ABI.
 *
 * Each entry point replaces one piece of the reference's command surface or
 * in-process API (paths relative to the reference root):
 *
 *   afx_device_count / afx_device_name   lsd, sd, sdt      anc-field/src/main.cpp:38-89
 *   afx_create                           Simulation ctor   anc-field/src/simulation.hpp:158-171
 *   afx_load_kerkythea                   load              main.cpp:97-100, simulation.hpp:173-199,
 *                                                          anc-field/src/kerkythea.hpp:63-196
 *   afx_set_min_size                     minsize           main.cpp:128-132, simulation.hpp:279-284
 *   afx_set_voxels / afx_set_materials   (Kerkythea output volume_data/coef_a/coef_b, kerkythea.hpp:29-31)
 *   afx_set_pml                          pml               main.cpp:134-137, simulation.hpp:233-277
 *   afx_add_source                       addsource         main.cpp:102-105, simulation.hpp:309-324
 *   afx_add_receiver                     addmic            main.cpp:107-110, simulation.hpp:326-340
 *   afx_run                              run               main.cpp:117-122, simulation.hpp:342-434
 *   afx_get_receivers                    save / mics()     main.cpp:124-127, simulation.hpp:95
 *   afx_get_pressure                     p()               simulation.hpp:94
 *   afx_get_plane                        vis cross section copy_plane.cl:29-34, simulation.hpp:286-289,394-401
 *   afx_add_room, afx_*_cell             (no counterpart: batched rooms and cell-addressed setup)
 *   afx_get_stats                        (boundary counts logged at lrs-data.hpp:212-214)
 *   afx_autotune                         (no counterpart: tile-shape and time-block search)
 *   afx_set_time_block, afx_get_time_block (no counterpart: two steps per pass)
 *   afx_set_tile_split, afx_get_tile_split (no counterpart: tiles dealt out in halves)
 *   afx_profile_run                      (no counterpart: per-kernel event timing)
 *   afx_synchronize                      clh Device::queue().finish(), simulation.hpp:407
 *
 * Conventions: every function returns AFX_OK (0) or a negative AFX_E_* code
 * and sets a thread-local message readable with afx_last_error().  No C++
 * exception crosses this boundary.  Callers own their arrays, which are
 * copied during the call; the library owns all device memory.  One handle is
 * used by one host thread at a time; different handles are independent.
 * All grids are x-fastest: index = (z*ny + y)*nx + x.
 */
#ifndef AFX_H
#define AFX_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AFX_OK 0
#define AFX_E_ARG (-1)       /* invalid argument */
#define AFX_E_STATE (-2)     /* call not valid in the handle's current state */
#define AFX_E_GEOMETRY (-3)  /* unphysical voxel pattern (lrs-data.hpp:327) */
#define AFX_E_IO (-4)        /* file missing or unreadable */
#define AFX_E_HIP (-5)       /* HIP runtime error */
#define AFX_E_NOMEM (-6)     /* allocation failed */
#define AFX_E_RANGE (-7)     /* position or size outside the grid / limits */

/* field storage precision (arithmetic is FP32 for FP32/FP16/BF16, FP64 for FP64) */
#define AFX_FP32 0
#define AFX_FP64 1
#define AFX_FP16 2
#define AFX_BF16 3

/* flags */
#define AFX_COMPAT 1u       /* reproduce the reference kernels' quirks (boundary gets lambda^2,
                               corner group X_M_Y_M_Z_M not updated, corners drop material 2,
                               corner ring stride ncoef/2, pml.cl over the whole interior) */
#define AFX_NO_GRAPH 2u     /* launch every step eagerly instead of replaying a captured hipGraph */
#define AFX_SPLIT 4u        /* separate interior and boundary launches (no fused step kernel) */
#define AFX_LIST_BLOCKS 8u  /* fused kernel: boundary and PML cells in workgroups of their own instead of
                               in the tile workgroups that hold them (same results) */
#define AFX_PML_BLOCKS 16u  /* fused kernel: physical PML stepped by z-marching blocks of PML columns
                               instead of one thread per PML cell record (same results; slower on
                               MI355X, whose latency the per-cell records hide better) */

/* PML faces */
#define AFX_FACE_XM 1u
#define AFX_FACE_XP 2u
#define AFX_FACE_YM 4u
#define AFX_FACE_YP 8u
#define AFX_FACE_ZM 16u
#define AFX_FACE_ZP 32u

typedef struct afx_sim afx_sim;

typedef struct afx_config {
    float c;            /* speed of sound, m/s          (reference 340)       */
    float dx;           /* cell size, m                 (reference 0.02)      */
    float dt;           /* time step, s                 (reference 3.125e-5)  */
    float res;          /* metres per cell used to quantise positions and voxelise models (0.02) */
    int32_t precision;  /* AFX_FP32 ... */
    uint32_t flags;     /* AFX_COMPAT | ... */
    int32_t device;     /* HIP device ordinal */
    float pml_order;    /* physical PML: polynomial grading order (default 3) */
    float pml_r0;       /* physical PML: design normal-incidence reflection (default 1e-3) */
    float scale;        /* FP16/BF16 storage: power-of-two field scale (0: the smallest power of two at or
                           above the peak of the sources added before the first afx_run; fixed from then on) */
} afx_config;

typedef struct afx_stats {
    int64_t cells;          /* sum over rooms of nx*ny*nz */
    int64_t boundary;       /* boundary cells over all rooms */
    int64_t faces, edges, corners;
    int64_t group_count[26];/* per reference group id (lrs-data.hpp:53-78) */
    int64_t steps;          /* steps run so far */
    int32_t rooms;
    int32_t nx, ny, nz;     /* room 0 */
    int32_t tile_x, tile_y, tile_z; /* interior tile in use */
    int32_t fused;          /* 1 when the fused step kernel is in use */
    double last_run_ms;     /* wall time of the last afx_run (device-synchronised) */
    double last_run_gpu_ms; /* device time of the last afx_run: HIP events on the handle's stream */
    int64_t last_run_steps; /* steps of the last afx_run */
} afx_stats;

/* configuration with the reference's constants: c=340, dx=0.02, dt=3.125e-5,
 * res=0.02 (simulation.hpp:162-166, main.cpp:64), FP32, AFX_COMPAT */
int afx_config_default(afx_config* cfg);

int afx_device_count(int32_t* n);
int afx_device_name(int32_t i, char* buf, size_t len);

int afx_create(const afx_config* cfg, afx_sim** out);
void afx_destroy(afx_sim* sim);
const char* afx_last_error(void);

/* Geometry of room 0.  afx_load_kerkythea voxelises axis-aligned boxes of a
 * Kerkythea XML scene at cfg.res and reads <material>_a.dat/_b.dat beside it. */
int afx_load_kerkythea(afx_sim* sim, const char* model_xml);
/* overwrites the grid dims with RoundTo((size_t)(x/res)+1, 32) per axis */
int afx_set_min_size(afx_sim* sim, float x, float y, float z);
int afx_set_voxels(afx_sim* sim, int32_t nx, int32_t ny, int32_t nz, const int32_t* vox /* -1 air */);
int afx_set_materials(afx_sim* sim, const float* a, const float* b, int32_t ncoef, int32_t nmat);

/* Absorbing layer.  Compat: profile[len] is placed reversed at both ends of
 * every axis like SetPml (NULL: all-zero, the shipped state whose profile
 * file is missing) and face_mask is ignored.  Physical: width cells on each
 * face in face_mask, graded zeta = zeta_max*(d/width)^order, or profile[len]
 * (len == width, zeta at depth 1..width) when profile is not NULL. */
int afx_set_pml(afx_sim* sim, int32_t width, uint32_t face_mask, const float* profile, int32_t len);

/* Sources add samples[t] to the cell's pressure after the boundary update of
 * step t (zero beyond n).  Positions in metres are quantised to (int)(pos/res). */
int afx_add_source(afx_sim* sim, const float pos[3], const float* samples, int64_t n);
int afx_add_receiver(afx_sim* sim, const float pos[3]);
int afx_add_source_cell(afx_sim* sim, int32_t room, const int32_t cell[3], const float* samples, int64_t n);
int afx_add_receiver_cell(afx_sim* sim, int32_t room, const int32_t cell[3]);

/* Batched identification: independent rooms stepped together.  All rooms share
 * c, dx, dt and ncoef.  Room 0 is the one set by afx_set_voxels/materials. */
int afx_add_room(afx_sim* sim, int32_t nx, int32_t ny, int32_t nz, const int32_t* vox,
                 const float* a, const float* b, int32_t ncoef, int32_t nmat, int32_t* room_id);

/* Advance every room by `steps` (resumable: a second call continues in time). */
int afx_run(afx_sim* sim, int64_t steps);
/* Host-only helpers (no device needed).
 * afx_voxelize_kerkythea: the voxel grid and material filters `load` would
 * build (kerkythea.hpp:133-196); NULL buffers only report dims/ncoef/nmat.
 * afx_classify: boundary cells per reference group (lrs-data.hpp:117-215);
 * AFX_E_GEOMETRY with the offending cell in bad_cell on an unphysical pattern. */
int afx_voxelize_kerkythea(const char* model_xml, float res, int32_t dims[3], int32_t* vox, int64_t vox_cap,
                           float* a, float* b, int32_t coef_cap, int32_t* ncoef, int32_t* nmat);
int afx_classify(int32_t nx, int32_t ny, int32_t nz, const int32_t* vox, int32_t compat, int64_t counts[26],
                 int32_t* bad_cell);

/* Waits until all work on the handle's device has finished (hipDeviceSynchronize). */
int afx_synchronize(afx_sim* sim);
/* Receiver traces [nmic][T] with T = steps run so far, in receiver order;
 * out == NULL only reports the count in *n_written. */
int afx_get_receivers(afx_sim* sim, float* out, int64_t cap, int64_t* n_written);
/* which: 0 = newest pressure p^{n+1}, 1 = p^n, 2 = p^{n-1}; FP32 copy of room `room` */
int afx_get_pressure(afx_sim* sim, int32_t room, int32_t which, float* out, int64_t cap);
/* One axis-aligned plane of a pressure field (which as in afx_get_pressure),
 * FP32: axis 0 gives [nz][ny] at x = index, axis 1 [nz][nx] at y = index
 * (the cross section copy_plane.cl:29-34 takes for the live view,
 * simulation.hpp:394-401), axis 2 [ny][nx] at z = index.  Only the plane
 * crosses PCIe. */
int afx_get_plane(afx_sim* sim, int32_t room, int32_t axis, int32_t index, int32_t which, float* out, int64_t cap);
int afx_get_stats(afx_sim* sim, afx_stats* st);

/* Per-kernel timing on the handle's stream: runs `steps` steps eagerly with a
 * HIP event pair around every launch and returns the average duration per
 * launch of each kernel class (the state advances like afx_run). */
typedef struct afx_profile {
    double step_ms;         /* average device time per step (first launch start to last launch end) */
    double kernel_ms[4];    /* average per launch: 0 step kernel (fused, or interior when split),
                               1 boundary / source / receiver, 2 PML pressure, 3 PML auxiliary */
    int64_t launches[4];
    int64_t steps;
} afx_profile;
int afx_profile_run(afx_sim* sim, int64_t steps, afx_profile* out);

/* Selects one of the compiled interior tile shapes (x, y, z cells per
 * workgroup): 64x4x16 (default), 64x4x32, 64x8x16, 64x2x32, 128x2x16, 64x4x8,
 * 256x1x16, and with 4 cells of x per lane 256x4x8 and 256x2x8, with 2 cells
 * 128x4x16 and 128x4x8 (these need every room width to be a multiple of that
 * count).  Results do not depend on the shape. */
int afx_set_tile(afx_sim* sim, int32_t tx, int32_t ty, int32_t tz);

/* Time-step blocking.  With steps = 2 each pass over the grid advances two
 * steps: every interior tile whose cells and one-cell halo are all plain
 * interior cells (no boundary, PML, source or receiver cell, none on the grid
 * edge) computes step n+1 on itself and its halo and step n+2 on itself in
 * one launch, reading p^n and p^{n-1} once (16 instead of 24 bytes of HBM
 * traffic per two FP32 cell updates); all other cells take the two steps in
 * two launches.  Results are bit-identical to steps = 1.  It takes a fourth
 * pressure field and applies to the fused step path (not with AFX_SPLIT,
 * AFX_LIST_BLOCKS, the compat PML, or one-row tile shapes, where steps stay
 * single).  Default 1. */
int afx_set_time_block(afx_sim* sim, int32_t steps);
/* The steps per pass set, and (pair_tiles, may be NULL) how many tiles take
 * two steps per pass on the committed handle (0 when single steps run). */
int afx_get_time_block(afx_sim* sim, int32_t* steps, int64_t* pair_tiles);
/* The fused step deals the tiles whose origin lies in the top quarter
 * (split 1) or the top half (split 2) of each room's planes as two work items
 * of half the tile's planes each; 0 (the default) keeps whole tiles.  The
 * step's last workgroups are then shorter, and so is the time the step waits
 * for them.  Results are bit-identical for every setting; afx_autotune
 * chooses it along with the tile shape. */
int afx_set_tile_split(afx_sim* sim, int32_t split);
int afx_get_tile_split(afx_sim* sim, int32_t* split);

/* Times the candidate interior tile shapes, each with one and two steps per
 * pass where that applies and with two assignments of tile layers to the
 * XCDs, then the fastest of them with tiles split (afx_set_tile_split 1 and
 * 2), on this handle's rooms (each is first checked bit-exact against the
 * current setting on a copy of the state; the best of three timed passes
 * counts), keeps the fastest and returns its tile in tile[3]
 * (afx_get_time_block gives its steps per pass, afx_get_tile_split its
 * split).  Leaves the simulation state unchanged.  Results are kept per
 * (device, precision, mode, PML, flags, room dims) for the process and in the
 * file AFX_TUNE_CACHE (default $XDG_CACHE_HOME/afx/tune-v2.txt or
 * $HOME/.cache/afx/tune-v2.txt;
 * "off" disables it), so a later process with the same key skips the
 * search. */
int afx_autotune(afx_sim* sim, int32_t steps_per_trial, int32_t tile[3]);

/* Measurement aid, no counterpart in the reference: a plain device copy of
 * `bytes` (read `bytes`, write `bytes`, float4 grid-stride kernel k_copy)
 * repeated `reps` times after one warm-up; *gbps = 2*bytes*reps / device time.
 * It is the calibration kernel for HBM counters and the practical copy rate
 * next to the 8 TB/s peak. */
int afx_copy_probe(int32_t device, int64_t bytes, int32_t reps, double* gbps);

#ifdef __cplusplus
}
#endif
#endif
