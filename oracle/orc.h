/*
 * CPU oracle for the FDTD room-acoustics step — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference algorithm (szarvas/anc-field) used
 * by tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py as
 * the checker.  Nothing in the product library links, loads or calls it.
 *
 * Function -> reference text it restates (paths relative to the reference
 * root) and how each is pinned:
 *   orc_classify     anc-field/src/lrs-data.hpp:117-215 (scan), :218-330
 *                    (pattern -> group), :24-78 (group ids); the scan stays
 *                    inside the voxel volume (:101-105,154-158).
 *                    Pinned: brute-force pattern enumeration and a numpy
 *                    brute force of the scan range (tests), and the
 *                    boundary-cell counts of examples/models/<name> (tests).
 *   leapfrog         anc-field/src/leapfrog.cl:34-55.
 *                    Pinned: discrete dispersion relation / rigid-box modes,
 *                    free-field arrival, the ie224 reference-held traces.
 *   lrs (26 groups)  lrs_1.cl:49-180, lrs_2.cl:49-216, lrs_3.cl:49-187,
 *                    host sequence simulation.hpp:358-370 (ring index :365,
 *                    lambda2 passed as `l` :368-369).
 *                    Pinned: ie224 reference-held traces (ncoef=1 faces,
 *                    edges, corners in compat mode); normal-incidence
 *                    reflection R=(1-beta)/(1+beta) for ncoef=1 (physical);
 *                    for ncoef 3 and 4, every face/edge/corner ring history
 *                    against scipy.signal.lfilter(b, a, x) and
 *                    y = (p+ - p-)/l, and the duct reflection spectrum of a
 *                    second-order wall against the scheme's exact discrete
 *                    reflection coefficient (tests/test_oracle_iir.py, with
 *                    planted mistakes that must fail).  Compat corners'
 *                    ring stride ncoef/2 (lrs_3.cl:70) is pinned only by
 *                    reading against the reference: parity unpinned there.
 *   pml (compat)     pml.cl:74-265, simulation.hpp:233-277, :373-382,
 *                    :411-430.  Pinned at the shipped state (kappa == 0,
 *                    where it must reduce to the plain leapfrog) and, with
 *                    the intended profile, by its absorption, the weak echo
 *                    of its graded inner face and the untouched direct
 *                    sound (tests).  The pressure update with non-zero
 *                    kappa, corner term kappa_x kappa_y kappa_z psi
 *                    included, is pinned by the decay rates of a uniform
 *                    field, which must converge to exp(-kappa dt) at third
 *                    order (tests/test_oracle_pml.py, with planted
 *                    mistakes).  Parity unpinned for the phi-update quirks
 *                    (pml.cl:124-137, :247), which enter only non-uniform
 *                    fields with non-zero kappa and are checked by reading.
 *   pml (physical)   no reference counterpart (Grote & Sim 2010 PML for the
 *                    scalar wave equation, the model pml.cl/sym_pml.py set
 *                    out to discretise).  Pinned: reduces to leapfrog when
 *                    zeta == 0; absorption against a rigid/Dirichlet wall.
 *   source / mic     source.cl:32-40, mic.cl:32-41, simulation.hpp:384-392,
 *                    positions simulation.hpp:322,338.
 *   storage rounding no reference counterpart (FP16/BF16 field storage).
 *                    orc_round_store is pinned against numpy's float16 and
 *                    torch's bfloat16 conversions over normal, subnormal,
 *                    tie and overflow inputs (tests).
 */
#ifndef AFX_ORACLE_H
#define AFX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* solid-face bits of a boundary cell, reference face order X_P..Z_M */
#define ORC_XP 1u
#define ORC_XM 2u
#define ORC_YP 4u
#define ORC_YM 8u
#define ORC_ZP 16u
#define ORC_ZM 32u

typedef struct {
    int32_t nx, ny, nz;
    const int32_t* vox;     /* [nz][ny][nx], -1 air, else material id; may be NULL (all air) */
    int32_t ncoef, nmat;
    const float* a;         /* [nmat][ncoef] admittance numerator  */
    const float* b;         /* [nmat][ncoef] admittance denominator */
    float c, dx, dt;        /* lambda = dt*c/dx */
    int32_t compat;         /* 1: reference quirks (boundary gets lambda^2, corner quirks) */
    /* sources and receivers (cell coordinates) */
    int32_t nsrc;
    const int32_t* src_cell;  /* [nsrc][3] */
    const float* src;         /* [nsrc][src_len] */
    int64_t src_len;
    int32_t nmic;
    const int32_t* mic_cell;  /* [nmic][3] */
    /* absorbing layer: 0 none, 1 compat pml.cl, 2 physical (Grote-Sim) */
    int32_t pml_mode;
    int32_t pml_width;
    const double* ksi[3];     /* per-axis profiles: compat: kappa[nx|ny|nz];
                                 physical: zeta at cell centres.  Double, so
                                 the f64 instantiation keeps the derived zeta
                                 unrounded; f32 rounds it once on load */
    const double* ksi_h[3];   /* physical: zeta at +1/2 points */
    const int32_t* pml_boxes; /* physical: [nbox][6] x0,x1,y0,y1,z0,z1 (half-open), disjoint */
    int32_t pml_nbox;
    /* reduced-precision field storage (f32 instantiation only): 0 none,
     * 1 IEEE binary16, 2 bfloat16.  Fields hold pressure / scale (a power of
     * two, <= 0 meaning 1); every pressure write rounds to the storage type,
     * sources enter as sample * (1/scale), traces are value * scale. */
    int32_t store;
    double scale;
    /* compat: size of the voxel volume the classifier scans (lrs-data.hpp:
     * 101-105), <= nx, ny, nz; 0 means the whole grid.  Cells of a grid grown
     * beyond it by minsize stay plain leapfrog cells. */
    int32_t vol[3];
} orc_problem;

/* Boundary classification restating lrs-data.hpp.  Writes up to `cap`
 * entries in the reference's order (group 0..25, scan order within a group):
 * pos[i*6 + {0..5}] = x, y, z, m0, m1, m2 (unused materials = -1), group[i].
 * Returns the number of boundary cells, or -1 on an unphysical pattern
 * (lrs-data.hpp:327), writing the offending cell into pos[0..2]. */
int64_t orc_classify(int32_t nx, int32_t ny, int32_t nz, const int32_t* vox,
                     int32_t* pos, int32_t* group, int64_t cap);
/* The same scan restricted to 1 <= x < lx-1, 1 <= y < ly-1, 1 <= z < lz-1
 * (the voxel volume's extent inside a larger grid). */
int64_t orc_classify_range(int32_t nx, int32_t ny, int32_t nz, const int32_t* vox,
                           int32_t lx, int32_t ly, int32_t lz,
                           int32_t* pos, int32_t* group, int64_t cap);

typedef struct orc_sim_f32 orc_sim_f32;
typedef struct orc_sim_f64 orc_sim_f64;

orc_sim_f32* orc_create_f32(const orc_problem* pb);
int          orc_run_f32(orc_sim_f32* s, int64_t steps, float* traces /*[nmic][steps] or NULL*/);
/* which: 0 = newest p^{n+1}, 1 = p^n, 2 = p^{n-1}; writes nx*ny*nz values */
void         orc_field_f32(const orc_sim_f32* s, int which, float* out);
/* overwrites one of the three pressure fields (which as above), e.g. to
 * start from a prescribed state */
void         orc_set_field_f32(orc_sim_f32* s, int which, const float* in);
void         orc_destroy_f32(orc_sim_f32* s);
int64_t      orc_nbound_f32(const orc_sim_f32* s);
/* IIR state of the boundary cells in classify order: [nb][3*ncoef] each */
void         orc_rings_f32(const orc_sim_f32* s, float* in, float* out);
/* ring slot ix_buf the last step wrote (simulation.hpp:229, :365) */
int32_t      orc_ring_index_f32(const orc_sim_f32* s);

orc_sim_f64* orc_create_f64(const orc_problem* pb);
int          orc_run_f64(orc_sim_f64* s, int64_t steps, double* traces);
void         orc_field_f64(const orc_sim_f64* s, int which, double* out);
void         orc_set_field_f64(orc_sim_f64* s, int which, const double* in);
void         orc_destroy_f64(orc_sim_f64* s);
int64_t      orc_nbound_f64(const orc_sim_f64* s);
void         orc_rings_f64(const orc_sim_f64* s, double* in, double* out);
int32_t      orc_ring_index_f64(const orc_sim_f64* s);

void         orc_set_threads(int n);

/* x rounded to nearest-even in the storage type `mode` (1 binary16,
 * 2 bfloat16) and widened back to float; mode 0 returns x. */
float        orc_round_store(float x, int32_t mode);

#ifdef __cplusplus
}
#endif
#endif
