/*
 * Oracle time stepping, instantiated twice by orc.c (REAL = float, double).
 * TEST INFRASTRUCTURE ONLY — see orc.h for the reference citations.
 *
 * Arithmetic is written in the operation order of the reference kernels so
 * that an FP32 build without contraction (-ffp-contract=off) reproduces the
 * reference's IEEE results operation for operation.
 */
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define SIM CAT(orc_sim, SFX)
#define FN(name) CAT(name, SFX)

struct SIM {
    orc_problem pb;
    int32_t nx, ny, nz;
    int64_t n;
    REAL* buf[3];        /* pressure storage; role permuted each step like simulation.hpp:360 */
    int rot[3];          /* rot[0] = output p_[0], rot[1] = p^n, rot[2] = p^{n-1} */
    /* boundary cells in classify order */
    int64_t nb;
    int32_t* bpos;       /* [nb][6] */
    int32_t* bgroup;     /* [nb] */
    REAL* ring_in;       /* [nb][3*ncoef] */
    REAL* ring_out;
    int32_t ix_buf;      /* simulation.hpp:229, :365 */
    int64_t sim_cnt;     /* simulation.hpp:432 */
    REAL lambda2, lam_b, c2i;   /* interior lambda^2, boundary 'l', 2(1-3 lambda^2) */
    REAL* a;             /* [nmat][ncoef] */
    REAL* b;
    /* sources / mics */
    REAL* src;
    /* pml */
    REAL* fi[3][2];      /* compat: fi_x/y/z pairs (new, old) ; physical: phi x/y/z in [d][0] */
    REAL* psi[2];
    int fi_rot;
    REAL* ksi[3];
    REAL* ksi_h[3];
    REAL* snap;          /* compat pml: unused scratch */
    int32_t store;       /* storage rounding mode (orc_problem.store) */
    REAL scale, inv_scale;
};

/* a pressure value as the field storage holds it */
static inline REAL FN(rnd)(const struct SIM* s, REAL v)
{
    return s->store ? (REAL)orc_round_store((float)v, s->store) : v;
}

static inline int64_t FN(ix)(const struct SIM* s, int64_t x, int64_t y, int64_t z)
{
    return (z * s->ny + y) * s->nx + x;
}

void FN(orc_destroy)(struct SIM* s);

struct SIM* FN(orc_create)(const orc_problem* pb)
{
    struct SIM* s = (struct SIM*)calloc(1, sizeof(struct SIM));
    s->pb = *pb;
    s->nx = pb->nx; s->ny = pb->ny; s->nz = pb->nz;
    s->n = (int64_t)pb->nx * pb->ny * pb->nz;
    for (int i = 0; i < 3; ++i) {
        s->buf[i] = (REAL*)calloc((size_t)s->n, sizeof(REAL));
        s->rot[i] = i;
    }
    /* lambda^2 exactly as simulation.hpp:166 */
    REAL c = (REAL)pb->c, dx = (REAL)pb->dx, dt = (REAL)pb->dt;
    s->lambda2 = (dt * c / dx) * (dt * c / dx);
    /* simulation.hpp:368-369 passes lambda2_ where the kernels expect lambda */
    s->lam_b = pb->compat ? s->lambda2 : (dt * c / dx);
    s->c2i = (REAL)2.0 * ((REAL)1.0 - (REAL)3.0 * s->lambda2);
    if (pb->store != 0 && sizeof(REAL) != sizeof(float)) { FN(orc_destroy)(s); return NULL; }
    s->store = pb->store;
    s->scale = (REAL)(pb->scale > 0 ? pb->scale : 1.0);
    s->inv_scale = (REAL)(1.0 / (pb->scale > 0 ? pb->scale : 1.0));
    int64_t nab = (int64_t)pb->nmat * pb->ncoef;
    s->a = (REAL*)calloc((size_t)(nab > 0 ? nab : 1), sizeof(REAL));
    s->b = (REAL*)calloc((size_t)(nab > 0 ? nab : 1), sizeof(REAL));
    for (int64_t i = 0; i < nab; ++i) { s->a[i] = (REAL)pb->a[i]; s->b[i] = (REAL)pb->b[i]; }
    s->nb = 0;
    if (pb->vox) {
        int64_t cap = s->n;
        int32_t* pos = (int32_t*)malloc((size_t)cap * 6 * sizeof(int32_t));
        int32_t* grp = (int32_t*)malloc((size_t)cap * sizeof(int32_t));
        const int32_t lx = pb->vol[0] > 0 ? pb->vol[0] : pb->nx;
        const int32_t ly = pb->vol[1] > 0 ? pb->vol[1] : pb->ny;
        const int32_t lz = pb->vol[2] > 0 ? pb->vol[2] : pb->nz;
        int64_t nb = orc_classify_range(pb->nx, pb->ny, pb->nz, pb->vox, lx, ly, lz, pos, grp, cap);
        if (nb < 0) { free(pos); free(grp); FN(orc_destroy)(s); return NULL; }
        s->nb = nb; s->bpos = pos; s->bgroup = grp;
        int64_t nr = nb * 3 * (pb->ncoef > 0 ? pb->ncoef : 1);
        s->ring_in = (REAL*)calloc((size_t)(nr > 0 ? nr : 1), sizeof(REAL));
        s->ring_out = (REAL*)calloc((size_t)(nr > 0 ? nr : 1), sizeof(REAL));
    }
    s->ix_buf = 0;
    s->sim_cnt = 0;
    if (pb->nsrc > 0) {
        s->src = (REAL*)malloc((size_t)(pb->nsrc * pb->src_len) * sizeof(REAL));
        for (int64_t i = 0; i < pb->nsrc * pb->src_len; ++i) s->src[i] = (REAL)pb->src[i];
    }
    if (pb->pml_mode != 0) {
        for (int d = 0; d < 3; ++d)
            for (int k = 0; k < 2; ++k) s->fi[d][k] = (REAL*)calloc((size_t)s->n, sizeof(REAL));
        s->psi[0] = (REAL*)calloc((size_t)s->n, sizeof(REAL));
        s->psi[1] = (REAL*)calloc((size_t)s->n, sizeof(REAL));
        int32_t dims[3] = {s->nx, s->ny, s->nz};
        int32_t dmax = s->nx > s->ny ? s->nx : s->ny;
        dmax = dmax > s->nz ? dmax : s->nz;
        for (int d = 0; d < 3; ++d) {
            /* sized for the largest axis: pml.cl:247 indexes ksi_x with y */
            s->ksi[d] = (REAL*)calloc((size_t)dmax + 2, sizeof(REAL));
            s->ksi_h[d] = (REAL*)calloc((size_t)dmax + 2, sizeof(REAL));
            for (int i = 0; i < dims[d]; ++i) {
                if (pb->ksi[d]) s->ksi[d][i] = (REAL)pb->ksi[d][i];
                if (pb->ksi_h[d]) s->ksi_h[d][i] = (REAL)pb->ksi_h[d][i];
            }
        }
        s->fi_rot = 0;
    }
    return s;
}

void FN(orc_destroy)(struct SIM* s)
{
    if (!s) return;
    for (int i = 0; i < 3; ++i) free(s->buf[i]);
    free(s->bpos); free(s->bgroup); free(s->ring_in); free(s->ring_out);
    free(s->a); free(s->b); free(s->src);
    for (int d = 0; d < 3; ++d) { free(s->fi[d][0]); free(s->fi[d][1]); free(s->ksi[d]); free(s->ksi_h[d]); }
    free(s->psi[0]); free(s->psi[1]); free(s->snap);
    free(s);
}

/* leapfrog.cl:34-55 over 0<x<nx-1 (likewise y, z), solid voxels included */
static void FN(leapfrog)(struct SIM* s, REAL* p, const REAL* p_d1, const REAL* p_d2)
{
    const int64_t nx = s->nx, ny = s->ny, nz = s->nz;
    const int64_t sy = nx, sz = nx * ny;
    const REAL l2 = s->lambda2, c2 = s->c2i;
    #pragma omp parallel for schedule(static)
    for (int64_t z = 1; z < nz - 1; ++z)
        for (int64_t y = 1; y < ny - 1; ++y) {
            const int64_t row = z * sz + y * sy;
            for (int64_t x = 1; x < nx - 1; ++x) {
                const int64_t i = row + x;
                REAL sum = p_d1[i + 1] + p_d1[i - 1];
                sum = sum + p_d1[i + sy];
                sum = sum + p_d1[i - sy];
                sum = sum + p_d1[i + sz];
                sum = sum + p_d1[i - sz];
                REAL v = l2 * sum + c2 * p_d1[i];
                p[i] = FN(rnd)(s, v - p_d2[i]);
            }
        }
}

/* Mirrored neighbour sum of lrs_1/2/3.cl: axes in x, y, z order; a solid
 * face replaces its pair by 2*opposite, a free axis contributes +1 then -1. */
static inline REAL FN(mirror_sum)(const struct SIM* s, const REAL* p, int64_t i, uint32_t m)
{
    const int64_t st[3] = {1, s->nx, (int64_t)s->nx * s->ny};
    REAL acc = 0;
    int first = 1;
    for (int d = 0; d < 3; ++d) {
        const uint32_t plus = 1u << (2 * d), minus = 2u << (2 * d);
        if (m & plus) {
            REAL t = (REAL)2.0 * p[i - st[d]];
            acc = first ? t : acc + t; first = 0;
        } else if (m & minus) {
            REAL t = (REAL)2.0 * p[i + st[d]];
            acc = first ? t : acc + t; first = 0;
        } else {
            acc = first ? p[i + st[d]] : acc + p[i + st[d]]; first = 0;
            acc = acc + p[i - st[d]];
        }
    }
    return acc;
}

static const uint32_t FN(group_mask)[26] = {
    ORC_XP, ORC_XM, ORC_YP, ORC_YM, ORC_ZP, ORC_ZM,
    ORC_XP | ORC_YP, ORC_XP | ORC_YM, ORC_XP | ORC_ZP, ORC_XP | ORC_ZM,
    ORC_XM | ORC_YP, ORC_XM | ORC_YM, ORC_XM | ORC_ZP, ORC_XM | ORC_ZM,
    ORC_YP | ORC_ZP, ORC_YP | ORC_ZM, ORC_YM | ORC_ZP, ORC_YM | ORC_ZM,
    ORC_XP | ORC_YP | ORC_ZP, ORC_XP | ORC_YP | ORC_ZM, ORC_XP | ORC_YM | ORC_ZP, ORC_XP | ORC_YM | ORC_ZM,
    ORC_XM | ORC_YP | ORC_ZP, ORC_XM | ORC_YP | ORC_ZM, ORC_XM | ORC_YM | ORC_ZP, ORC_XM | ORC_YM | ORC_ZM,
};

/* One boundary cell: lrs_1.cl:116-179 (one material), lrs_2.cl:62-215 (two),
 * lrs_3.cl:62-186 (three). */
static void FN(lrs_cell)(struct SIM* s, int64_t e, REAL* p_p1, const REAL* p, const REAL* p_d1)
{
    const int32_t* q = s->bpos + e * 6;
    const int grp = s->bgroup[e];
    const uint32_t m = FN(group_mask)[grp];
    const int K = grp <= 5 ? 1 : (grp <= 17 ? 2 : 3);
    const int N = s->pb.ncoef;
    const int compat = s->pb.compat;
    const int64_t i = FN(ix)(s, q[0], q[1], q[2]);
    REAL* rin = s->ring_in + e * 3 * N;
    REAL* rout = s->ring_out + e * 3 * N;
    const REAL l = s->lam_b;
    const unsigned ixb = (unsigned)s->ix_buf;
    /* ring stride: N, except compat corners use size_a_0/2 (lrs_3.cl:70) */
    const unsigned sf = (compat && K == 3) ? (unsigned)N / 2u : (unsigned)N;
    REAL a0[3], b0[3], g[3];
    for (int k = 0; k < K; ++k) {
        const int mat = q[3 + k];
        const REAL* ak = s->a + (int64_t)mat * N;
        const REAL* bk = s->b + (int64_t)mat * N;
        a0[k] = ak[0]; b0[k] = bk[0];
        REAL gn = 0;
        for (unsigned ii = 1; ii < (unsigned)N; ++ii) {
            const unsigned slot = (ixb + ii) % sf + (unsigned)k * sf;
            gn += bk[ii] * rin[slot] - ak[ii] * rout[slot];
        }
        g[k] = gn;
    }
    const REAL l2 = l * l;
    const REAL cb = (REAL)2.0 * ((REAL)1.0 - (REAL)3.0 * l2);
    const int update_p = !(compat && grp == 25);  /* simulation.hpp:452 drops "X_M_Y_M" */
    if (update_p) {
        const REAL S = FN(mirror_sum)(s, p, i, m);
        REAL num, den;
        if (K == 1) {
            num = l2 * S + cb * p[i];
            num = num + l2 / b0[0] * g[0];
            num = num + (l * a0[0] / b0[0] - (REAL)1.0) * p_d1[i];
            den = (REAL)1.0 + l * a0[0] / b0[0];
        } else {
            const int U = (compat && K == 3) ? 2 : K;  /* lrs_3.cl:96-174 use materials 0 and 1 */
            REAL G = g[0] / b0[0];
            REAL A = l * a0[0] / b0[0];
            den = (REAL)1.0 + l * a0[0] / b0[0];
            for (int k = 1; k < U; ++k) {
                G = G + g[k] / b0[k];
                A = A + l * a0[k] / b0[k];
                den = den + l * a0[k] / b0[k];
            }
            num = l2 * S + l2 * G;
            num = num + cb * p[i];
            num = num + (A - (REAL)1.0) * p_d1[i];
        }
        p_p1[i] = FN(rnd)(s, num / den);
    }
    const REAL pp = p_p1[i];
    for (int k = 0; k < K; ++k) {
        const REAL xn = a0[k] / (l * b0[k]) * (pp - p_d1[i]) - g[k] / b0[k];
        const unsigned w = ixb + (unsigned)k * sf;
        rin[w] = xn;
        rout[w] = (REAL)1.0 / a0[k] * (b0[k] * xn + g[k]);
    }
}

/* ---- compat PML: pml.cl:154-265 with simulation.hpp:373-382 ---- */
static inline REAL FN(avg4)(const REAL* f, int64_t i0, int64_t da, int64_t db)
{
    return (REAL)0.25 * (f[i0] + f[i0 + db] + f[i0 + da] + f[i0 + da + db]);
}

static void FN(pml_compat)(struct SIM* s, REAL* p, const REAL* pm1, const REAL* pm2)
{
    const int64_t nx = s->nx, ny = s->ny, nz = s->nz, sy = nx, sz = nx * ny;
    const int L = s->pb.pml_width;
    const REAL c = (REAL)s->pb.c, dx = (REAL)s->pb.dx, dt = (REAL)s->pb.dt;
    /* std::rotate on two-element vectors swaps new/old each step */
    s->fi_rot ^= 1;
    const int nw = s->fi_rot, od = s->fi_rot ^ 1;
    REAL *fx = s->fi[0][nw], *fxm1 = s->fi[0][od];
    REAL *fy = s->fi[1][nw], *fym1 = s->fi[1][od];
    REAL *fz = s->fi[2][nw], *fzm1 = s->fi[2][od];
    REAL *psiph = s->psi[nw], *psimh = s->psi[od];
    const REAL *kx = s->ksi[0], *ky = s->ksi[1], *kz = s->ksi[2];
    /* The reference kernel reads p and psiph at +1 neighbours while other
     * work-items write them.  Defined here as three ordered phases: psi, p,
     * then fi (each reading the completed previous phase). */
    #pragma omp parallel for schedule(static)
    for (int64_t z = 2; z < nz - 2; ++z)
        for (int64_t y = 2; y < ny - 2; ++y)
            for (int64_t x = 2; x < nx - 2; ++x) {
                const int64_t i = z * sz + y * sy + x;
                psiph[i] = dt * pm1[i] + psimh[i];
            }
    #pragma omp parallel for schedule(static)
    for (int64_t z = 2; z < nz - 2; ++z)
        for (int64_t y = 2; y < ny - 2; ++y)
            for (int64_t x = 2; x < nx - 2; ++x) {
                const int64_t i = z * sz + y * sy + x;
                REAL lap = pm1[i + 1] + pm1[i - 1] + pm1[i + sy] + pm1[i - sy] + pm1[i + sz] + pm1[i - sz]
                           - (REAL)6.0 * pm1[i];
                REAL dv = FN(avg4)(fxm1, i + 1, sy, sz) - FN(avg4)(fxm1, i, sy, sz)
                          + FN(avg4)(fym1, i + sy, 1, sz) - FN(avg4)(fym1, i, 1, sz)
                          + FN(avg4)(fzm1, i + sz, 1, sy) - FN(avg4)(fzm1, i, 1, sy);
                REAL t1 = c * c / (dx * dx) * lap;
                REAL t2 = (REAL)1.0 / dx * dv;
                REAL t3 = kx[x] * ky[y] * kz[z] * (REAL)0.5 * (psiph[i] + psimh[i]);
                REAL t4 = (REAL)1.0 / (dt * dt) * (pm2[i] - (REAL)2.0 * pm1[i]);
                REAL t5 = (kx[x] * ky[y] + ky[y] * kz[z] + kz[z] * kx[x]) * pm1[i];
                REAL t6 = (kx[x] + ky[y] + kz[z]) / ((REAL)2.0 * dt) * pm2[i];
                REAL num = t1 + t2 - t3 - t4 - t5 + t6;
                p[i] = FN(rnd)(s, num / ((REAL)1.0 / (dt * dt) + (kx[x] + ky[y] + kz[z]) / ((REAL)2.0 * dt)));
            }
    #pragma omp parallel for schedule(static)
    for (int64_t z = 2; z < nz - 2; ++z)
        for (int64_t y = 2; y < ny - 2; ++y)
            for (int64_t x = 2; x < nx - 2; ++x) {
                const int64_t i = z * sz + y * sy + x;
                const REAL kax = (REAL)0.5 * (kx[x] + kx[x + 1]);
                const REAL kay = (REAL)0.5 * (ky[y] + ky[y + 1]);
                const REAL kaz = (REAL)0.5 * (kz[z] + kz[z + 1]);
                if (x < L || x > nx - L - 1) {
                    /* Dxp: second average takes p where pm1 is meant (pml.cl:124-127) */
                    REAL d = (REAL)0.5 / dx * (FN(avg4)(p, i + 1, sy, sz) - FN(avg4)(p, i, sy, sz)
                                               + FN(avg4)(pm1, i + 1, sy, sz) - FN(avg4)(p, i, sy, sz));
                    REAL dp = (REAL)1.0 / dx * (FN(avg4)(psiph, i + 1, sy, sz) - FN(avg4)(psiph, i, sy, sz));
                    fx[i] = (-(REAL)0.5 * kax * fxm1[i] + (kay + kaz - kax) * d + (kay * kaz) * dp
                             + (REAL)1.0 / dt * (fxm1[i])) / ((REAL)1.0 / dt + kax * (REAL)0.5);
                }
                if (y < L || y > ny - L - 1) {
                    REAL d = (REAL)0.5 / dx * (FN(avg4)(p, i + sy, 1, sz) - FN(avg4)(p, i, 1, sz)
                                               + FN(avg4)(pm1, i + sy, 1, sz) - FN(avg4)(p, i, 1, sz));
                    REAL dp = (REAL)1.0 / dx * (FN(avg4)(psiph, i + sy, 1, sz) - FN(avg4)(psiph, i, 1, sz));
                    /* ksi_avg(ksi_x, y) as written at pml.cl:247 */
                    const REAL kxy = (REAL)0.5 * (kx[y] + kx[y + 1]);
                    fy[i] = (-(REAL)0.5 * kay * fym1[i] + (kax + kaz - kay) * d + (kxy * kaz) * dp
                             + (REAL)1.0 / dt * (fym1[i])) / ((REAL)1.0 / dt + kay * (REAL)0.5);
                }
                if (z < L || z > nz - L - 1) {
                    REAL d = (REAL)0.5 / dx * (FN(avg4)(p, i + sz, 1, sy) - FN(avg4)(p, i, 1, sy)
                                               + FN(avg4)(pm1, i + sz, 1, sy) - FN(avg4)(p, i, 1, sy));
                    REAL dp = (REAL)1.0 / dx * (FN(avg4)(psiph, i + sz, 1, sy) - FN(avg4)(psiph, i, 1, sy));
                    fz[i] = (-(REAL)0.5 * kaz * fzm1[i] + (kax + kay - kaz) * d + (kax * kay) * dp
                             + (REAL)1.0 / dt * (fzm1[i])) / ((REAL)1.0 / dt + kaz * (REAL)0.5);
                }
            }
}

/* ---- physical PML (Grote-Sim), pressure part: runs before the boundary cells ---- */
static void FN(pml_phys_p)(struct SIM* s, REAL* p1, const REAL* p, const REAL* pm)
{
    const int64_t nx = s->nx, ny = s->ny, sy = nx, sz = nx * ny;
    const REAL dx = (REAL)s->pb.dx, dt = (REAL)s->pb.dt;
    const REAL l2 = s->lambda2;
    REAL *phx = s->fi[0][0], *phy = s->fi[1][0], *phz = s->fi[2][0], *psi = s->psi[0];
    const REAL *zx = s->ksi[0], *zy = s->ksi[1], *zz = s->ksi[2];
    for (int bI = 0; bI < s->pb.pml_nbox; ++bI) {
        const int32_t* bx = s->pb.pml_boxes + 6 * bI;
        #pragma omp parallel for schedule(static)
        for (int64_t z = bx[4]; z < bx[5]; ++z)
            for (int64_t y = bx[2]; y < bx[3]; ++y)
                for (int64_t x = bx[0]; x < bx[1]; ++x) {
                    const int64_t i = z * sz + y * sy + x;
                    const REAL sx = zx[x], syy = zy[y], szz = zz[z];
                    const REAL s1 = sx + syy + szz;
                    const REAL s2 = sx * syy + syy * szz + szz * sx;
                    const REAL s3 = sx * syy * szz;
                    REAL sum = p[i + 1] + p[i - 1];
                    sum = sum + p[i + sy];
                    sum = sum + p[i - sy];
                    sum = sum + p[i + sz];
                    sum = sum + p[i - sz];
                    const REAL dv = (phx[i] - phx[i - 1]) + (phy[i] - phy[i - sy]) + (phz[i] - phz[i - sz]);
                    const REAL psn = psi[i] + dt * p[i];
                    REAL num = l2 * (sum - (REAL)6.0 * p[i]);
                    num = num + (dt * dt / dx) * dv;
                    num = num - (dt * dt) * s3 * (REAL)0.5 * (psn + psi[i]);
                    num = num - (dt * dt) * s2 * p[i];
                    num = num + (REAL)2.0 * p[i] - pm[i];
                    num = num + (REAL)0.5 * dt * s1 * pm[i];
                    p1[i] = FN(rnd)(s, num / ((REAL)1.0 + (REAL)0.5 * dt * s1));
                    psi[i] = psn;
                }
    }
}

/* ---- physical PML, auxiliary phi: runs after the step's pressure is final ---- */
static void FN(pml_phys_phi)(struct SIM* s, const REAL* p1, const REAL* p)
{
    const int64_t nx = s->nx, ny = s->ny, sy = nx, sz = nx * ny;
    const REAL c = (REAL)s->pb.c, dx = (REAL)s->pb.dx, dt = (REAL)s->pb.dt;
    REAL *phx = s->fi[0][0], *phy = s->fi[1][0], *phz = s->fi[2][0], *psi = s->psi[0];
    const REAL *zx = s->ksi[0], *zy = s->ksi[1], *zz = s->ksi[2];
    const REAL *hx = s->ksi_h[0], *hy = s->ksi_h[1], *hz = s->ksi_h[2];
    const REAL cc = c * c, idt = (REAL)1.0 / dt, idx = (REAL)1.0 / dx;
    for (int bI = 0; bI < s->pb.pml_nbox; ++bI) {
        const int32_t* bx = s->pb.pml_boxes + 6 * bI;
        #pragma omp parallel for schedule(static)
        for (int64_t z = bx[4]; z < bx[5]; ++z)
            for (int64_t y = bx[2]; y < bx[3]; ++y)
                for (int64_t x = bx[0]; x < bx[1]; ++x) {
                    const int64_t i = z * sz + y * sy + x;
                    {
                        const REAL h = hx[x];
                        const REAL g2 = zy[y] + zz[z] - h, g3 = zy[y] * zz[z];
                        const REAL dn = (p1[i + 1] - p1[i]) * idx, dold = (p[i + 1] - p[i]) * idx;
                        const REAL dps = (psi[i + 1] - psi[i]) * idx;
                        phx[i] = ((idt - (REAL)0.5 * h) * phx[i] + cc * g2 * (REAL)0.5 * (dn + dold) + cc * g3 * dps)
                                 / (idt + (REAL)0.5 * h);
                    }
                    {
                        const REAL h = hy[y];
                        const REAL g2 = zx[x] + zz[z] - h, g3 = zx[x] * zz[z];
                        const REAL dn = (p1[i + sy] - p1[i]) * idx, dold = (p[i + sy] - p[i]) * idx;
                        const REAL dps = (psi[i + sy] - psi[i]) * idx;
                        phy[i] = ((idt - (REAL)0.5 * h) * phy[i] + cc * g2 * (REAL)0.5 * (dn + dold) + cc * g3 * dps)
                                 / (idt + (REAL)0.5 * h);
                    }
                    {
                        const REAL h = hz[z];
                        const REAL g2 = zx[x] + zy[y] - h, g3 = zx[x] * zy[y];
                        const REAL dn = (p1[i + sz] - p1[i]) * idx, dold = (p[i + sz] - p[i]) * idx;
                        const REAL dps = (psi[i + sz] - psi[i]) * idx;
                        phz[i] = ((idt - (REAL)0.5 * h) * phz[i] + cc * g2 * (REAL)0.5 * (dn + dold) + cc * g3 * dps)
                                 / (idt + (REAL)0.5 * h);
                    }
                }
    }
}

/* simulation.hpp:342-434 */
int FN(orc_run)(struct SIM* s, int64_t steps, REAL* traces)
{
    const orc_problem* pb = &s->pb;
    const int N = pb->ncoef;
    for (int64_t it = 0; it < steps; ++it) {
        /* std::rotate(p_.begin(), p_.begin()+2, p_.end()) */
        int r0 = s->rot[0], r1 = s->rot[1], r2 = s->rot[2];
        s->rot[0] = r2; s->rot[1] = r0; s->rot[2] = r1;
        REAL* p_p1 = s->buf[s->rot[0]];
        REAL* p = s->buf[s->rot[1]];
        REAL* p_d1 = s->buf[s->rot[2]];
        FN(leapfrog)(s, p_p1, p, p_d1);
        if (pb->pml_mode == 2) FN(pml_phys_p)(s, p_p1, p, p_d1);
        if (pb->vox) {
            if (s->ix_buf == 0) s->ix_buf = N - 1; else s->ix_buf -= 1;
            for (int64_t e = 0; e < s->nb; ++e) FN(lrs_cell)(s, e, p_p1, p, p_d1);
        }
        if (pb->pml_mode == 1) FN(pml_compat)(s, p_p1, p, p_d1);
        const int64_t t = s->sim_cnt + it;
        for (int i = 0; i < pb->nsrc; ++i) {
            const int32_t* q = pb->src_cell + 3 * i;
            REAL v = t < pb->src_len ? s->src[i * pb->src_len + t] : (REAL)0.0;
            p_p1[FN(ix)(s, q[0], q[1], q[2])] += v * s->inv_scale;
        }
        /* all of a cell's sources are added before it is stored */
        if (s->store)
            for (int i = 0; i < pb->nsrc; ++i) {
                const int32_t* q = pb->src_cell + 3 * i;
                const int64_t j = FN(ix)(s, q[0], q[1], q[2]);
                p_p1[j] = FN(rnd)(s, p_p1[j]);
            }
        if (traces)
            for (int i = 0; i < pb->nmic; ++i) {
                const int32_t* q = pb->mic_cell + 3 * i;
                traces[(int64_t)i * steps + it] = p_p1[FN(ix)(s, q[0], q[1], q[2])] * s->scale;
            }
        if (pb->pml_mode == 2) FN(pml_phys_phi)(s, p_p1, p);
        if (pb->pml_mode == 1 && it % 8000 == 0) {
            for (int d = 0; d < 3; ++d) {
                memset(s->fi[d][0], 0, (size_t)s->n * sizeof(REAL));
                memset(s->fi[d][1], 0, (size_t)s->n * sizeof(REAL));
            }
            memset(s->psi[0], 0, (size_t)s->n * sizeof(REAL));
            memset(s->psi[1], 0, (size_t)s->n * sizeof(REAL));
        }
    }
    s->sim_cnt += steps;
    return 0;
}

void FN(orc_field)(const struct SIM* s, int which, REAL* out)
{
    memcpy(out, s->buf[s->rot[which]], (size_t)s->n * sizeof(REAL));
}

void FN(orc_set_field)(struct SIM* s, int which, const REAL* in)
{
    memcpy(s->buf[s->rot[which]], in, (size_t)s->n * sizeof(REAL));
}

int64_t FN(orc_nbound)(const struct SIM* s) { return s->nb; }

int32_t FN(orc_ring_index)(const struct SIM* s) { return s->ix_buf; }

void FN(orc_rings)(const struct SIM* s, REAL* in, REAL* out)
{
    int64_t nr = s->nb * 3 * s->pb.ncoef;
    memcpy(in, s->ring_in, (size_t)nr * sizeof(REAL));
    memcpy(out, s->ring_out, (size_t)nr * sizeof(REAL));
}

#undef SIM
#undef FN
