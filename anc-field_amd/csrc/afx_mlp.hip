// Surrogate MLP: BF16 MFMA training and inference kernels and the
// afx_mlp_* C ABI (include/afx_mlp.h).
//
// No counterpart in the reference.  The MLP maps (source, receiver)
// positions in one room to the band-limited impulse response that
// anc_field_py/ancutil.py:49-79 otherwise obtains from a full FDTD run.
//
// Kernels (gfx950).  A GEMM workgroup is 4 waves on one 32x32 output tile:
// each wave takes every 4th 32-wide slice of K on 2x2
// v_mfma_f32_16x16x32_bf16 blocks with FP32 accumulators, and the four
// partial tiles are summed through LDS before the epilogue (the batch-256
// GEMMs have too few tiles to fill the chip one wave per tile):
//   k_features       Fourier features of the positions -> X [rows][F] and X^T
//   k_gemm<EPI>      C = A . B^T over K (A [M][K], B [N][K], both K-contiguous)
//                    with a fused epilogue:
//                      HIDDEN  z = C + b; keep z (FP32); a = silu(z) -> BF16 row- and column-major
//                      OUT     y = C + b (FP32)
//                      LOSS    y = C + b; g = 2 w (y - t) -> BF16 both layouts; per-wave loss partial
//                      BACK    g = C * silu'(z_prev) -> BF16 both layouts
//   k_gemm128<EPI>   HIDDEN and OUT for inference chunks (K a multiple of 128):
//                    128x128 tiles through LDS, k-slices visited in turn (same
//                    bits as k_gemm)
//   k_gemm64<EPI>    HIDDEN and OUT for inference chunks: 64x64 tiles through LDS,
//                    the same bits as k_gemm
//   k_grads          all weight gradients of a step, dW_l = dY_l^T . X_l (A = dY_l^T
//                    [out][B], B = X_l^T [in][B], contraction over the batch), and all
//                    bias gradients, db_l[n] = sum_b dY_l[b][n], in one launch, each
//                    with AdamW fused when training
//   k_adam           AdamW on the FP32 master weights, 32x32 tiles of W per
//                    workgroup; refreshes the BF16 copies W [out][in] (forward
//                    operand) and W^T [in][out] (backward operand, written
//                    through an LDS transpose)
//
// Every operand of an MFMA is read K-contiguous: the forward uses X and W,
// the input gradient dY and W^T, the weight gradient dY^T and X^T.  The
// epilogues that produce X and dY write both layouts, so no kernel transposes
// through LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <numeric>
#include <random>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/afx.h"
#include "../../include/afx_mlp.h"
#include "afx_launch.hpp"

namespace afx {
int report_error(int code, const char* msg);
}

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMaxLayers = 16;

enum Epi { EPI_HIDDEN = 0, EPI_OUT = 1, EPI_LOSS = 2, EPI_BACK = 3 };

// Training-step position kept on the device, so that a captured graph of
// steps replays unchanged: the minibatch rows, the loss partials and the
// learning-rate and bias-correction factors of step *step are found from
// it.  With step == null every offset is zero and the arguments are used as
// given.
struct StepRef {
    const int64_t* step;   // device counter, advanced once per run of steps
    int64_t rows_stride;   // minibatch rows advance per step
    int64_t part_stride;   // loss partials advance per step
    const float* sched;    // [step][3]: lr, 1 - beta1^t, 1 - beta2^t
    int64_t rel;           // this launch's step after the counter's value
};

__device__ inline int64_t step_of(const StepRef& r) { return (r.step ? *r.step : 0) + r.rel; }

struct Gemm {
    const bf16* A;       // [M][lda], K-contiguous
    const bf16* B;       // [N][ldb], K-contiguous
    int lda, ldb, M, N, K;
    const float* bias;   // [N]: HIDDEN, OUT, LOSS
    float* f32;          // HIDDEN: z, OUT/LOSS: y (either may be null in LOSS)
    int ldf;
    bf16* act;           // HIDDEN: silu(z); LOSS/BACK: gradient.  [M][ldact]
    bf16* actT;          // same values, [N][ldactT] (may be null)
    int ldact, ldactT;
    const float* zprev;  // BACK: pre-activation at this output, [M][ldz]
    int ldz;
    const float* target; // LOSS: [*][ldt], row rows[m] (or m when rows is null)
    int ldt;
    const int* rows;
    const float* wrow;   // LOSS: per-row weight
    float* partial;      // LOSS: one loss partial per workgroup
    StepRef sr;          // LOSS: rows and partial advance with the device step
};

struct GradProblem {
    const bf16* A;       // dY^T [M=out][lda]
    const bf16* B;       // X^T  [N=in][ldb]
    float* out;          // dW [M][N] (without the fused update)
    int lda, ldb, M, N, K;
    int64_t woff;        // parameter offset of W (fused update)
    bf16* wb;            // BF16 copies W [M][N] and W^T [N][M] (fused update)
    bf16* wt;
};

// AdamW applied where a gradient is produced (training): the gradient never
// goes to memory.
struct AdamArgs {
    float *w, *m1, *m2;
    float lr, b1, b2, eps, wd, bc1, bc2;
    int on;              // 0: write gradients instead
};

struct GradTable {
    GradProblem p[kMaxLayers];
    int start[kMaxLayers + 1];  // first workgroup of each problem
    int n;
};

struct ColsumTable {
    const bf16* g[kMaxLayers];  // dY^T [rows_l][ld]
    float* out[kMaxLayers];
    int64_t boff[kMaxLayers];   // parameter offset of b_l (fused update)
    int rows[kMaxLayers];
    int start[kMaxLayers + 1];  // first output row of each layer in the flattened range
    int ld, K, n;
};

struct AdamTable {
    int64_t off[kMaxLayers + 1];  // parameter offset of W_l; b_l follows W_l
    int rows[kMaxLayers], cols[kMaxLayers];
    bf16* wb[kMaxLayers];         // [rows][cols]
    bf16* wt[kMaxLayers];         // [cols][rows]
    int tstart[kMaxLayers + 1];   // first workgroup of layer l's 32x32 tiles of W
    int bofs[kMaxLayers + 1];     // first bias of layer l in the concatenated biases
    int n;
};

// SiLU and its derivative from the hardware exponential and reciprocal
// (v_exp_f32 on z log2 e, v_rcp_f32): relative error about 1e-6, far below
// the BF16 rounding of the stored activation (2^-8), and a few instructions
// instead of a libm exponential and an IEEE division.  Training and inference
// call the same functions, so their activations keep equal bits.
__device__ inline float sigmoid_fast(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ inline float silu(float z) { return z * sigmoid_fast(z); }

__device__ inline float silu_grad(float z)
{
    const float s = sigmoid_fast(z);
    return s + z * s * (1.0f - s);
}

// One wave computes the 32x32 block at (m0, n0) of A . B^T.  MFMA operand
// layout (16x16x32): lane l supplies row (l & 15) of A and row (l & 15) of
// B^T's transpose, 8 consecutive k at offset 8*(l >> 4); the result block has
// rows 4*(l >> 4) + e, e < 4, at column (l & 15).  A and B take the same k
// assignment, so the order of k within a lane does not matter.
__device__ inline void mma_block(const bf16* __restrict__ A, int lda, const bf16* __restrict__ B, int ldb, int K,
                                 int m0, int n0, f32x4 (&c)[2][2])
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int r = lane & 15, kq = (lane >> 4) << 3;
    const bf16* pa0 = A + (size_t)(m0 + r) * lda + kq;
    const bf16* pa1 = pa0 + (size_t)16 * lda;
    const bf16* pb0 = B + (size_t)(n0 + r) * ldb + kq;
    const bf16* pb1 = pb0 + (size_t)16 * ldb;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) c[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 4
    for (int k = 32 * wave; k < K; k += 128) {
        const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(pa0 + k);
        const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(pa1 + k);
        const bf16x8 b0 = *reinterpret_cast<const bf16x8*>(pb0 + k);
        const bf16x8 b1 = *reinterpret_cast<const bf16x8*>(pb1 + k);
        c[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, c[0][0], 0, 0, 0);
        c[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, c[0][1], 0, 0, 0);
        c[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, c[1][0], 0, 0, 0);
        c[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, c[1][1], 0, 0, 0);
    }
}

// AdamW on parameter i with gradient gi; returns its new value.
__device__ inline float adam_step(float* __restrict__ w, float* __restrict__ m1, float* __restrict__ m2, int64_t i,
                                  float gi, float lr, float b1, float b2, float eps, float wd, float bc1, float bc2)
{
    float v = w[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float s = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = s;
    const float mh = a / bc1, vh = s / bc2;
    v = v - lr * (mh / (sqrtf(vh) + eps) + wd * v);
    w[i] = v;
    return v;
}

// Learning rate and bias corrections of the current step.
__device__ inline void adam_sched(AdamArgs& a, const StepRef& sr)
{
    if (sr.sched) {
        const float* q = sr.sched + 3 * step_of(sr);
        a.lr = q[0];
        a.bc1 = q[1];
        a.bc2 = q[2];
    }
}

// Sums the four waves' partial 32x32 tiles through LDS.  Afterwards thread t
// holds, for block (i, j) = (q >> 1, q & 1) with q = t / 64 and lane
// l = t % 64, the four elements e < 4 at row 16 i + 4 (l >> 4) + e, column
// 16 j + (l & 15) of the tile (the layout of one wave's accumulators).
__device__ inline f32x4 reduce_waves(const f32x4 (&c)[2][2], float (*lds)[4][4][64])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int q = 0; q < 4; ++q)
        for (int e = 0; e < 4; ++e) lds[wave][q][e][lane] = c[q >> 1][q & 1][e];
    __syncthreads();
    const int q = wave;
    f32x4 r;
    for (int e = 0; e < 4; ++e) r[e] = ((lds[0][q][e][lane] + lds[1][q][e][lane]) + lds[2][q][e][lane]) + lds[3][q][e][lane];
    return r;
}

template <int E>
__global__ __launch_bounds__(256) void k_gemm(Gemm g)
{
    __shared__ float lds[4][4][4][64];
    __shared__ float lpart[4];
    const int tiles_n = g.N >> 5;
    const int m0 = (int)(blockIdx.x / tiles_n) << 5;
    const int n0 = (int)(blockIdx.x % tiles_n) << 5;
    f32x4 c[2][2];
    mma_block(g.A, g.lda, g.B, g.ldb, g.K, m0, n0, c);
    const f32x4 acc = reduce_waves(c, lds);

    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int i = q >> 1, j = q & 1;
    const int n = n0 + 16 * j + (lane & 15);
    const int mb = m0 + 16 * i + ((lane >> 4) << 2);
    float loss = 0.0f;
    bf16x4 col;
    // loaded once: the stores below may alias it as far as the compiler knows
    const float bn = E == EPI_BACK ? 0.0f : g.bias[n];
    for (int e = 0; e < 4; ++e) {
        const int m = mb + e;
        float v = acc[e];
        if (E == EPI_HIDDEN) {
            const float z = v + bn;
            if (g.f32) g.f32[(size_t)m * g.ldf + n] = z;
            v = silu(z);
        } else if (E == EPI_OUT) {
            g.f32[(size_t)m * g.ldf + n] = v + bn;
            continue;
        } else if (E == EPI_LOSS) {
            const float y = v + bn;
            if (g.f32) g.f32[(size_t)m * g.ldf + n] = y;
            const int src = g.rows ? g.rows[step_of(g.sr) * g.sr.rows_stride + m] : m;
            const float w = g.wrow[src];
            const float d = y - g.target[(size_t)src * g.ldt + n];
            loss += w * (d * d);
            v = 2.0f * w * d;
        } else {  // EPI_BACK
            v = v * silu_grad(g.zprev[(size_t)m * g.ldz + n]);
        }
        const bf16 h = (bf16)v;
        g.act[(size_t)m * g.ldact + n] = h;
        col[e] = h;
    }
    if (E != EPI_OUT && g.actT) *reinterpret_cast<bf16x4*>(g.actT + (size_t)n * g.ldactT + mb) = col;
    if (E == EPI_LOSS) {
        for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o);
        if (lane == 0) lpart[q] = loss;
        __syncthreads();
        if (threadIdx.x == 0)
            g.partial[step_of(g.sr) * g.sr.part_stride + blockIdx.x] = ((lpart[0] + lpart[1]) + lpart[2]) + lpart[3];
    }
}

// Logical output tile of this workgroup.  Workgroup b runs on XCD b % 8;
// when the grid is a multiple of 8, XCD x takes the contiguous run of tiles
// [x * nb/8, (x+1) * nb/8), in order, so the tiles of one row panel of A
// (consecutive in the row-major tile order) run together on one XCD and
// read that panel through its L2 once, instead of once per XCD.
__device__ __forceinline__ int xcd_tile()
{
    const int b = (int)blockIdx.x, nb = (int)gridDim.x;
    return (nb & 7) ? b : (b & 7) * (nb >> 3) + (b >> 3);
}

// Inference GEMM for large M (HIDDEN and OUT epilogues): a 64x64 output tile
// per workgroup, operands staged through LDS in k-steps of 32 (double
// buffered, one barrier per step), each wave a 32x32 quadrant.  It gives the
// same bits as k_gemm: every MFMA sees the same operands per lane, and each
// output keeps four partial sums, one per k-slice (k-step mod 4) as k_gemm's
// waves do, added in k_gemm's order.  Needs M % 64 == 0, N % 64 == 0,
// K % 128 == 0.
template <int E>
__global__ __launch_bounds__(256) void k_gemm64(Gemm g)
{
    constexpr int kPad = 40;  // LDS row stride in bf16 (32 + 8)
    __shared__ bf16 As[2][64][kPad];
    __shared__ bf16 Bs[2][64][kPad];
    const int tiles_n = g.N >> 6;
    const int tile = xcd_tile();
    const int m0 = (tile / tiles_n) << 6;
    const int n0 = (tile % tiles_n) << 6;
    const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = wave >> 1, wj = wave & 1;
    const int lr = tid >> 2, lc = (tid & 3) << 3;  // this thread's 16-byte piece of each k-step
    const bf16* ga = g.A + (size_t)(m0 + lr) * g.lda + lc;
    const bf16* gb = g.B + (size_t)(n0 + lr) * g.ldb + lc;
    bf16x8 ra = *reinterpret_cast<const bf16x8*>(ga);
    bf16x8 rb = *reinterpret_cast<const bf16x8*>(gb);
    *reinterpret_cast<bf16x8*>(&As[0][lr][lc]) = ra;
    *reinterpret_cast<bf16x8*>(&Bs[0][lr][lc]) = rb;
    __syncthreads();
    const int r = lane & 15, kq = (lane >> 4) << 3;
    f32x4 p[4][2][2];
    for (int sl = 0; sl < 4; ++sl)
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) p[sl][i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const int nk = g.K >> 5;
    for (int kt0 = 0; kt0 < nk; kt0 += 4) {
#pragma unroll
        for (int sl = 0; sl < 4; ++sl) {
            const int kt = kt0 + sl, cur = kt & 1;
            const bool more = kt + 1 < nk;
            if (more) {
                ra = *reinterpret_cast<const bf16x8*>(ga + 32 * (kt + 1));
                rb = *reinterpret_cast<const bf16x8*>(gb + 32 * (kt + 1));
            }
            const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(&As[cur][32 * wi + r][kq]);
            const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(&As[cur][32 * wi + 16 + r][kq]);
            const bf16x8 b0 = *reinterpret_cast<const bf16x8*>(&Bs[cur][32 * wj + r][kq]);
            const bf16x8 b1 = *reinterpret_cast<const bf16x8*>(&Bs[cur][32 * wj + 16 + r][kq]);
            p[sl][0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, p[sl][0][0], 0, 0, 0);
            p[sl][0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, p[sl][0][1], 0, 0, 0);
            p[sl][1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, p[sl][1][0], 0, 0, 0);
            p[sl][1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, p[sl][1][1], 0, 0, 0);
            if (more) {
                *reinterpret_cast<bf16x8*>(&As[cur ^ 1][lr][lc]) = ra;
                *reinterpret_cast<bf16x8*>(&Bs[cur ^ 1][lr][lc]) = rb;
            }
            __syncthreads();
        }
    }
    // biases loaded before any store; 32-bit offsets (the host checks
    // M x ld < 2^31 for this path)
    float bj[2];
    for (int j = 0; j < 2; ++j) bj[j] = g.bias[n0 + 32 * wj + 16 * j + r];
    const bool keep = g.f32 != nullptr;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            const int n = n0 + 32 * wj + 16 * j + r;
            const int mb = m0 + 32 * wi + 16 * i + ((lane >> 4) << 2);
            for (int e = 0; e < 4; ++e) {
                const float v = ((p[0][i][j][e] + p[1][i][j][e]) + p[2][i][j][e]) + p[3][i][j][e];
                const int m = mb + e;
                if (E == EPI_HIDDEN) {
                    const float z = v + bj[j];
                    if (keep) g.f32[m * g.ldf + n] = z;
                    g.act[m * g.ldact + n] = (bf16)silu(z);
                } else {
                    g.f32[m * g.ldf + n] = v + bj[j];
                }
            }
        }
}

// Inference GEMM for large M on 128x128 output tiles (HIDDEN and OUT
// epilogues): eight waves, each a 64x32 part of 4x2 MFMA blocks.  Operands
// go straight from memory into LDS (global_load_lds_dwordx4: 16 B per lane,
// no VGPRs), one k-step of 32 per stage, four stages in a ring: the loads
// of k-step v + 2 are issued while k-step v is computed, one barrier per
// k-step (a stage is refilled two k-steps after its last read, across a
// barrier).  Six stages (one workgroup per CU) measured 40 % slower.  LDS rows are 64 B, with the 16-byte chunk c of
// row r stored at chunk c ^ ((r >> 1) & 3): the fragment reads (16 rows per
// quarter wave) and the loads' LDS writes meet no bank conflicts.  The bits
// are k_gemm's: every MFMA sees the same operands per lane, and each output
// is the sum of four k-slice partials (slice = k-step mod 4) in k_gemm's
// order.  The k-steps are visited slice by slice (0, 4, 8, ..., then 1, 5,
// 9, ...), so only the current slice's partials and the running sum of the
// finished slices are held: ((p0 + p1) + p2) + p3 as in k_gemm64.  Needs M
// and N multiples of 128 and K of 128.
constexpr int kG128Stages = 4;  // 64 KiB of LDS: two workgroups per CU

// s_waitcnt with vmcnt(n) and the other counters left open (gfx9 encoding)
#define AFX_VM_WAIT(n) __builtin_amdgcn_s_waitcnt(0x0F70 | ((n) & 15) | ((((n) >> 4) & 3) << 14))

// LDS position (in bf16) of 16-byte chunk c of operand row `row`
__device__ __forceinline__ int g128_at(int row, int c) { return row * 32 + ((c ^ ((row >> 1) & 3)) << 3); }

// Six 16-byte fragment reads from LDS (four of A at byte offsets 0, 1024,
// 2048 and 3072 from pa, two of B at 0 and 1024 from pb), with the wait for
// them in the same block, so every output holds its value when the block
// ends.  Written out because the compiler, seeing LDS reads while loads into
// LDS are in flight, would wait for all of them first.
__device__ __forceinline__ void g128_frags(uint32_t pa, uint32_t pb, bf16x8 (&a)[4], bf16x8 (&b)[2])
{
    asm volatile(
        "ds_read_b128 %0, %6\n\t"
        "ds_read_b128 %1, %6 offset:1024\n\t"
        "ds_read_b128 %2, %6 offset:2048\n\t"
        "ds_read_b128 %3, %6 offset:3072\n\t"
        "ds_read_b128 %4, %7\n\t"
        "ds_read_b128 %5, %7 offset:1024\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(a[0]), "=&v"(a[1]), "=&v"(a[2]), "=&v"(a[3]), "=&v"(b[0]), "=&v"(b[1])
        : "v"(pa), "v"(pb)
        : "memory");
}

template <int E>
__global__ __launch_bounds__(512) void k_gemm128(Gemm g)
{
    constexpr int S = kG128Stages;
    static_assert(S >= 3 && 2 * (S - 2) < 64, "stages");
    __shared__ bf16 As[S][4096];  // [stage][row * 32 + k], chunks swizzled (g128_at)
    __shared__ bf16 Bs[S][4096];
    const int tiles_n = g.N >> 7;
    const int tile = xcd_tile();
    const int m0 = (tile / tiles_n) << 7;
    const int n0 = (tile % tiles_n) << 7;
    const int tid = (int)threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wi = wave >> 2, wj = wave & 3;
    const int nk = g.K >> 5, ns = nk >> 2;
    auto kof = [&](int v) { return 32 * (v / ns + 4 * (v % ns)); };
    // this lane's 16-byte chunk of a stage: slot q = tid holds row q >> 2,
    // stored chunk q & 3, i.e. logical chunk (q & 3) ^ ((row >> 1) & 3)
    const int lrow = tid >> 2;
    const int lchunk = (tid & 3) ^ ((lrow >> 1) & 3);
    const bf16* ga = g.A + (size_t)(m0 + lrow) * g.lda + 8 * lchunk;
    const bf16* gb = g.B + (size_t)(n0 + lrow) * g.ldb + 8 * lchunk;
    auto issue = [&](int v) {  // k-step v into stage v % S; past the end: the last k-step again
        const int k = kof(v < nk ? v : nk - 1);
        __builtin_amdgcn_global_load_lds(ga + k, &As[v % S][0] + wave * 512, 16, 0, 0);
        __builtin_amdgcn_global_load_lds(gb + k, &Bs[v % S][0] + wave * 512, 16, 0, 0);
    };
#pragma unroll
    for (int v = 0; v < S - 2; ++v) issue(v);
    const int r = lane & 15, kc = lane >> 4;
    // fragment rows 64 wi + 16 i + r (A) and 32 wj + 16 j + r (B) share the
    // chunk swizzle ((row >> 1) & 3 does not see the 16 i), so the four A
    // (two B) fragments sit 1024 bytes apart
    const uint32_t oa = 2u * (uint32_t)g128_at(64 * wi + r, kc), ob = 2u * (uint32_t)g128_at(32 * wj + r, kc);
    f32x4 sum[4][2], p[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            sum[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            p[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        }
    for (int v = 0; v < nk; ++v) {
        // two loads per wave per k-step, retired in order: once at most
        // 2 (S - 2) are outstanding, k-step v is in LDS
        issue(v + S - 2);
        AFX_VM_WAIT(2 * (S - 2));
        __builtin_amdgcn_s_barrier();
        bf16x8 a[4], b[2];
        g128_frags((uint32_t)(uintptr_t)&As[v % S][0] + oa, (uint32_t)(uintptr_t)&Bs[v % S][0] + ob, a, b);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                p[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], p[i][j], 0, 0, 0);
        if (v % ns == ns - 1) {  // slice v / ns complete
            const bool first = v < ns;
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    sum[i][j] = first ? p[i][j] : sum[i][j] + p[i][j];
                    p[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                }
        }
    }
    AFX_VM_WAIT(0);  // no load into LDS outlives the workgroup
    // biases loaded before any store (which the compiler would otherwise
    // assume may change them); 32-bit offsets (the host checks M x ld < 2^31)
    float bj[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) bj[j] = g.bias[n0 + 32 * wj + 16 * j + r];
    const bool keep = g.f32 != nullptr;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int n = n0 + 32 * wj + 16 * j + r;
            const int mb = m0 + 64 * wi + 16 * i + ((lane >> 4) << 2);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float v = sum[i][j][e];
                const int m = mb + e;
                if (E == EPI_HIDDEN) {
                    const float z = v + bj[j];
                    if (keep) g.f32[m * g.ldf + n] = z;
                    g.act[m * g.ldact + n] = (bf16)silu(z);
                } else {
                    g.f32[m * g.ldf + n] = v + bj[j];
                }
            }
        }
}

// Weight-gradient tile `bid` of the grouped problems (with AdamW fused when
// ad.on).
__device__ inline void grads_tile(const GradTable& t, AdamArgs ad, const StepRef& sr, int bid,
                                  float (*lds)[4][4][64])
{
    int p = 0;
    while (p + 1 < t.n && bid >= t.start[p + 1]) ++p;
    const GradProblem& g = t.p[p];
    const int tile = bid - t.start[p];
    const int tiles_n = g.N >> 5;
    const int m0 = (tile / tiles_n) << 5, n0 = (tile % tiles_n) << 5;
    f32x4 c[2][2];
    mma_block(g.A, g.lda, g.B, g.ldb, g.K, m0, n0, c);
    const f32x4 acc = reduce_waves(c, lds);
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int n = n0 + 16 * (q & 1) + (lane & 15);
    const int mb = m0 + 16 * (q >> 1) + ((lane >> 4) << 2);
    if (!ad.on) {
        for (int e = 0; e < 4; ++e) g.out[(size_t)(mb + e) * g.N + n] = acc[e];
        return;
    }
    adam_sched(ad, sr);
    bf16x4 col;
    for (int e = 0; e < 4; ++e) {
        const int64_t k = (int64_t)(mb + e) * g.N + n;
        const float v = adam_step(ad.w, ad.m1, ad.m2, g.woff + k, acc[e], ad.lr, ad.b1, ad.b2, ad.eps, ad.wd, ad.bc1,
                                  ad.bc2);
        col[e] = (bf16)v;
        g.wb[k] = col[e];
    }
    *reinterpret_cast<bf16x4*>(g.wt + (size_t)n * g.M + mb) = col;
}

// Fourier features: a block holds blockDim.z rows; thread (q, j, r) writes
// sin and cos of 2 pi f_k x[b][j] for the KV frequencies k = KV q .. KV q +
// KV - 1, b = blockIdx.x * blockDim.z + r, at features j*2F + k and
// j*2F + F + k, in X [rows][nfeat] (one KV-element store each; KV divides
// F) and X^T [nfeat][ldt].
template <int KV> struct alignas(2 * KV) Bf16Pack {
    bf16 v[KV];
};

template <int KV>
__global__ void k_features(const float* __restrict__ x, const int* __restrict__ rows, int nrows, int in_dim, int nf,
                           const float* __restrict__ freq, bf16* __restrict__ X, bf16* __restrict__ XT, int ldt,
                           StepRef sr)
{
    const int k0 = KV * (int)threadIdx.x, j = (int)threadIdx.y, b = (int)(blockIdx.x * blockDim.z + threadIdx.z);
    if (b >= nrows || k0 >= nf || j >= in_dim) return;
    const int src = rows ? rows[step_of(sr) * sr.rows_stride + b] : b;
    const float xj = x[(size_t)src * in_dim + j];
    Bf16Pack<KV> sv, cv;
#pragma unroll
    for (int u = 0; u < KV; ++u) {
        const float cyc = freq[k0 + u] * xj;
        // v_sin_f32 / v_cos_f32 take their argument in turns; cyc less its
        // nearest integer (exact in FP32) is the same angle within half a
        // turn, where they are accurate to about 1e-6, far below the BF16
        // rounding
        const float ph = cyc - rintf(cyc);
        sv.v[u] = (bf16)__builtin_amdgcn_sinf(ph);
        cv.v[u] = (bf16)__builtin_amdgcn_cosf(ph);
    }
    const int nfeat = 2 * in_dim * nf;
    const int fs = j * 2 * nf + k0, fc = fs + nf;
    *reinterpret_cast<Bf16Pack<KV>*>(X + (size_t)b * nfeat + fs) = sv;
    *reinterpret_cast<Bf16Pack<KV>*>(X + (size_t)b * nfeat + fc) = cv;
    if (XT) {
#pragma unroll
        for (int u = 0; u < KV; ++u) {
            XT[(size_t)(fs + u) * ldt + b] = sv.v[u];
            XT[(size_t)(fc + u) * ldt + b] = cv.v[u];
        }
    }
}

// Bias gradients of workgroup `bid`: one wave per output row of the
// flattened layer range (with AdamW fused when ad.on).
__device__ inline void colsum_rows(const ColsumTable& t, AdamArgs ad, const StepRef& sr, int bid)
{
    const int row = bid * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= t.start[t.n]) return;
    int l = 0;
    while (l + 1 < t.n && row >= t.start[l + 1]) ++l;
    const int n = row - t.start[l];
    const bf16* g = t.g[l] + (size_t)n * t.ld;
    float s = 0.0f;
    for (int b = lane; b < t.K; b += 64) s += (float)g[b];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane != 0) return;
    if (!ad.on) {
        t.out[l][n] = s;
        return;
    }
    adam_sched(ad, sr);
    adam_step(ad.w, ad.m1, ad.m2, t.boff[l] + n, s, ad.lr, ad.b1, ad.b2, ad.eps, ad.wd, ad.bc1, ad.bc2);
}

// All weight gradients (workgroups [0, G), G = gt.start[gt.n]) and all bias
// gradients (the rest) of one step in one launch; the two parts are
// independent.
__global__ __launch_bounds__(256) void k_grads(GradTable gt, ColsumTable ct, AdamArgs ad, StepRef sr)
{
    __shared__ float lds[4][4][4][64];
    const int bid = (int)blockIdx.x, ng = gt.start[gt.n];
    if (bid < ng)
        grads_tile(gt, ad, sr, bid, lds);
    else
        colsum_rows(ct, ad, sr, bid - ng);
}

// AdamW on parameter i; returns its new value (the old one when !update).
__device__ inline float adam_param(float* __restrict__ w, float* __restrict__ m1, float* __restrict__ m2,
                                   const float* __restrict__ grad, int64_t i, float lr, float b1, float b2, float eps,
                                   float wd, float bc1, float bc2, int update)
{
    float v = w[i];
    if (update) {
        const float gi = grad[i];
        const float a = b1 * m1[i] + (1.0f - b1) * gi;
        const float s = b2 * m2[i] + (1.0f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = s;
        const float mh = a / bc1, vh = s / bc2;
        v = v - lr * (mh / (sqrtf(vh) + eps) + wd * v);
        w[i] = v;
    }
    return v;
}

// AdamW over all parameters (update == 0 only refreshes the BF16 copies).
// Workgroups [0, tstart[n]) take 32x32 tiles of the weight matrices: each
// thread updates 4 consecutive weights of a row, writes their BF16 copies to
// W and, through an LDS transpose, 4 consecutive elements of a row of W^T.
// The remaining workgroups take the biases, 256 each.
__global__ __launch_bounds__(256) void k_adam(AdamTable t, float* __restrict__ w, float* __restrict__ m1,
                                              float* __restrict__ m2, const float* __restrict__ grad, float lr,
                                              float b1, float b2, float eps, float wd, float bc1, float bc2, int update,
                                              StepRef sr)
{
    __shared__ bf16 tr[32][34];
    if (sr.sched) {
        const float* q = sr.sched + 3 * step_of(sr);
        lr = q[0];
        bc1 = q[1];
        bc2 = q[2];
    }
    const int bid = (int)blockIdx.x;
    if (bid >= t.tstart[t.n]) {
        const int k = (bid - t.tstart[t.n]) * 256 + (int)threadIdx.x;
        if (k >= t.bofs[t.n]) return;
        int l = 0;
        while (l + 1 < t.n && k >= t.bofs[l + 1]) ++l;
        const int64_t i = t.off[l] + (int64_t)t.rows[l] * t.cols[l] + (k - t.bofs[l]);
        adam_param(w, m1, m2, grad, i, lr, b1, b2, eps, wd, bc1, bc2, update);
        return;
    }
    int l = 0;
    while (l + 1 < t.n && bid >= t.tstart[l + 1]) ++l;
    const int tile = bid - t.tstart[l];
    const int cols = t.cols[l], rows = t.rows[l];
    const int r0 = (tile / (cols >> 5)) << 5, c0 = (tile % (cols >> 5)) << 5;
    const int rr = (int)threadIdx.x >> 3, cc = ((int)threadIdx.x & 7) << 2;
    bf16x4 h;
    for (int e = 0; e < 4; ++e) {
        const int64_t i = t.off[l] + (int64_t)(r0 + rr) * cols + c0 + cc + e;
        h[e] = (bf16)adam_param(w, m1, m2, grad, i, lr, b1, b2, eps, wd, bc1, bc2, update);
        tr[cc + e][rr] = h[e];
    }
    *reinterpret_cast<bf16x4*>(t.wb[l] + (size_t)(r0 + rr) * cols + c0 + cc) = h;
    __syncthreads();
    bf16x4 o;
    for (int e = 0; e < 4; ++e) o[e] = tr[rr][cc + e];
    *reinterpret_cast<bf16x4*>(t.wt[l] + (size_t)(c0 + rr) * rows + r0 + cc) = o;
}

__global__ void k_step_inc(int64_t* step, int64_t k)
{
    if (threadIdx.x == 0) *step += k;
}

__global__ void k_step_zero(int64_t* step)
{
    if (threadIdx.x == 0) *step = 0;
}

#define HIP_CHECK(expr)                                                                                  \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess) {                                                                          \
            char b_[512];                                                                                \
            snprintf(b_, sizeof b_, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
            return afx::report_error(AFX_E_HIP, b_);                                                     \
        }                                                                                                \
    } while (0)

int err(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int err(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return afx::report_error(code, buf);
}

struct Arena {
    char* base = nullptr;
    size_t used = 0, cap = 0;
    template <typename T>
    T* take(size_t n)
    {
        const size_t bytes = (n * sizeof(T) + 255) & ~size_t(255);
        T* p = reinterpret_cast<T*>(base + used);
        used += bytes;
        return p;
    }
};

}  // namespace

// Device memory of afx_mlp_train_noalloc: minibatch indices, learning-rate
// schedule, loss partials and the step counter of one call, laid out as the
// buffers afx_mlp_train allocates per call.  Each GPU holds its own copy
// until the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns on that device (scratch_setup).
static __device__ __align__(256) unsigned char scratch[AFX_MLP_TRAIN_SCRATCH_BYTES];

struct afx_mlp {
    afx_mlp_config cfg{};
    int device = 0;
    int L = 0, F = 0, B = 0;
    std::vector<int> in, out;       // per layer
    std::vector<int64_t> off;       // parameter offsets (W_l, b_l follows), size L+1
    int64_t nparam = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    void* mem = nullptr;
    // device buffers
    float *w = nullptr, *m1 = nullptr, *m2 = nullptr, *g = nullptr, *freq = nullptr;
    std::vector<bf16*> wb, wt;
    bf16 *X = nullptr, *XT = nullptr;
    std::vector<float*> Z;          // hidden pre-activations [B][out_l]
    std::vector<bf16*> A, AT;       // hidden activations
    std::vector<bf16*> D, DT;       // output gradients of each layer
    float* Y = nullptr;             // [B][out]
    float *xq = nullptr, *tq = nullptr, *wq = nullptr;  // staging for forward/gradients
    float* partial = nullptr;       // loss partials of one step
    // inference: chunks of up to kInferRows rows, own buffers (allocated on
    // first use) so a forward pass is a few large GEMMs instead of many
    // batch-sized ones
    void* imem = nullptr;
    bf16* Xi = nullptr;
    bf16* Ai[2] = {nullptr, nullptr};  // hidden activations, ping-pong
    float* Yi = nullptr;
    float* xqi = nullptr;
    StepRef sr{};                   // device step position while training (null otherwise)
    int nblk_loss = 0;
    // training set
    float *xd = nullptr, *td = nullptr, *wd = nullptr;
    int64_t ndata = 0;
    std::vector<int> perm;          // host permutation of the training rows (sized by set_data)
    unsigned char* scratch_dev = nullptr;  // this device's `scratch` (device address)
    int64_t t = 0;                  // optimizer steps taken
    afx_mlp_stats st{};
    std::vector<float> freq_host;
};

namespace {

int gemm_launch(afx_mlp* s, int epi, Gemm g)
{
    if (g.M % 32 || g.N % 32 || g.K % 32) return err(AFX_E_ARG, "gemm %dx%dx%d not a multiple of 32", g.M, g.N, g.K);
    // large-M inference GEMMs on 128x128 or 64x64 tiles (same bits; see
    // k_gemm128 and k_gemm64), whose epilogues index outputs in 32 bits
    const bool off32 = (int64_t)g.M * std::max(g.ldf, g.ldact) < ((int64_t)1 << 31);
    if ((epi == EPI_HIDDEN || epi == EPI_OUT) && !g.actT && off32 && g.M >= 1024 && g.M % 128 == 0 &&
        g.N % 128 == 0 && g.K % 128 == 0) {
        const dim3 grid128((unsigned)((g.M / 128) * (g.N / 128))), block128(512);
        HIP_CHECK(afx::launch_ok(grid128, block128, 512));
        if (epi == EPI_HIDDEN)
            k_gemm128<EPI_HIDDEN><<<grid128, block128, 0, s->stream>>>(g);
        else
            k_gemm128<EPI_OUT><<<grid128, block128, 0, s->stream>>>(g);
        HIP_CHECK(hipGetLastError());
        return AFX_OK;
    }
    if ((epi == EPI_HIDDEN || epi == EPI_OUT) && !g.actT && off32 && g.M >= 1024 && g.M % 64 == 0 &&
        g.N % 64 == 0 && g.K % 128 == 0) {
        const dim3 grid64((unsigned)((g.M / 64) * (g.N / 64))), block64(256);
        HIP_CHECK(afx::launch_ok(grid64, block64, 256));
        if (epi == EPI_HIDDEN)
            k_gemm64<EPI_HIDDEN><<<grid64, block64, 0, s->stream>>>(g);
        else
            k_gemm64<EPI_OUT><<<grid64, block64, 0, s->stream>>>(g);
        HIP_CHECK(hipGetLastError());
        return AFX_OK;
    }
    const dim3 grid((unsigned)((g.M / 32) * (g.N / 32))), block(256);
    HIP_CHECK(afx::launch_ok(grid, block, 256));
    switch (epi) {
    case EPI_HIDDEN: k_gemm<EPI_HIDDEN><<<grid, block, 0, s->stream>>>(g); break;
    case EPI_OUT: k_gemm<EPI_OUT><<<grid, block, 0, s->stream>>>(g); break;
    case EPI_LOSS: k_gemm<EPI_LOSS><<<grid, block, 0, s->stream>>>(g); break;
    default: k_gemm<EPI_BACK><<<grid, block, 0, s->stream>>>(g); break;
    }
    HIP_CHECK(hipGetLastError());
    return AFX_OK;
}

// Features of `nrows` positions (gathered through rows when not null), then
// the hidden layers.  With keep set, pre-activations and the column-major
// copies needed by the backward pass are written.
hipError_t launch_features(afx_mlp* s, const float* x, const int* rows, int nrows, bf16* X, bf16* XT, int ldt,
                           StepRef sr)
{
    // nf / KV x in_dim threads per row (at most 1024, checked at creation),
    // rows stacked in z up to about 256 threads per block; four frequencies
    // per thread (8-byte stores) when n_freq allows
    const int kv = s->cfg.n_freq % 4 == 0 ? 4 : 1;
    const int per_row = s->cfg.n_freq / kv * s->cfg.in_dim;
    const int rz = std::max(1, std::min(256 / per_row, 1024 / per_row));
    const dim3 block((unsigned)(s->cfg.n_freq / kv), (unsigned)s->cfg.in_dim, (unsigned)rz);
    const dim3 grid((unsigned)((nrows + rz - 1) / rz));
    if (hipError_t le = afx::launch_ok(grid, block, 1024)) return le;
    if (kv == 4)
        k_features<4><<<grid, block, 0, s->stream>>>(x, rows, nrows, s->cfg.in_dim, s->cfg.n_freq, s->freq, X, XT, ldt,
                                                     sr);
    else
        k_features<1><<<grid, block, 0, s->stream>>>(x, rows, nrows, s->cfg.in_dim, s->cfg.n_freq, s->freq, X, XT, ldt,
                                                     sr);
    return hipGetLastError();
}

int forward_hidden(afx_mlp* s, const float* x, const int* rows, bool keep)
{
    HIP_CHECK(launch_features(s, x, rows, s->B, s->X, keep ? s->XT : nullptr, s->B, s->sr));
    for (int l = 0; l + 1 < s->L; ++l) {
        Gemm g{};
        g.A = l == 0 ? s->X : s->A[l - 1];
        g.lda = s->in[l];
        g.B = s->wb[l];
        g.ldb = s->in[l];
        g.M = s->B;
        g.N = s->out[l];
        g.K = s->in[l];
        g.bias = s->w + s->off[l] + (int64_t)s->out[l] * s->in[l];
        g.f32 = keep ? s->Z[l] : nullptr;
        g.ldf = s->out[l];
        g.act = s->A[l];
        g.ldact = s->out[l];
        g.actT = keep ? s->AT[l] : nullptr;
        g.ldactT = s->B;
        int rc = gemm_launch(s, EPI_HIDDEN, g);
        if (rc) return rc;
    }
    return AFX_OK;
}

// One forward + backward over the batch in X (rows gathered from x/target/weight).
int step_grad(afx_mlp* s, const float* x, const float* target, const float* weight, const int* rows, float* y,
              bool fused_adam = false)
{
    int rc = forward_hidden(s, x, rows, true);
    if (rc) return rc;
    const int L = s->L;
    {
        Gemm g{};
        g.A = L >= 2 ? s->A[L - 2] : s->X;
        g.lda = s->in[L - 1];
        g.B = s->wb[L - 1];
        g.ldb = s->in[L - 1];
        g.M = s->B;
        g.N = s->out[L - 1];
        g.K = s->in[L - 1];
        g.bias = s->w + s->off[L - 1] + (int64_t)s->out[L - 1] * s->in[L - 1];
        g.f32 = y;
        g.ldf = s->out[L - 1];
        g.act = s->D[L - 1];
        g.ldact = s->out[L - 1];
        g.actT = s->DT[L - 1];
        g.ldactT = s->B;
        g.target = target;
        g.ldt = s->out[L - 1];
        g.rows = rows;
        g.wrow = weight;
        g.partial = s->partial;
        g.sr = s->sr;
        if ((rc = gemm_launch(s, EPI_LOSS, g))) return rc;
    }
    for (int l = L - 1; l >= 1; --l) {  // input gradient of layer l = output gradient of layer l-1
        Gemm g{};
        g.A = s->D[l];
        g.lda = s->out[l];
        g.B = s->wt[l];
        g.ldb = s->out[l];
        g.M = s->B;
        g.N = s->in[l];
        g.K = s->out[l];
        g.act = s->D[l - 1];
        g.ldact = s->in[l];
        g.actT = s->DT[l - 1];
        g.ldactT = s->B;
        g.zprev = s->Z[l - 1];
        g.ldz = s->in[l];
        if ((rc = gemm_launch(s, EPI_BACK, g))) return rc;
    }
    GradTable gt{};
    gt.n = L;
    gt.start[0] = 0;
    for (int l = 0; l < L; ++l) {
        GradProblem& p = gt.p[l];
        p.A = s->DT[l];
        p.lda = s->B;
        p.B = l == 0 ? s->XT : s->AT[l - 1];
        p.ldb = s->B;
        p.out = s->g + s->off[l];
        p.M = s->out[l];
        p.N = s->in[l];
        p.K = s->B;
        p.woff = s->off[l];
        p.wb = s->wb[l];
        p.wt = s->wt[l];
        gt.start[l + 1] = gt.start[l] + (p.M / 32) * (p.N / 32);
    }
    AdamArgs ad{};
    if (fused_adam) {
        s->t += 1;
        ad = AdamArgs{s->w, s->m1, s->m2, 0.0f, s->cfg.beta1, s->cfg.beta2, s->cfg.eps, s->cfg.weight_decay,
                      1.0f, 1.0f, 1};
    }
    ColsumTable ct{};
    ct.n = L;
    ct.ld = s->B;
    ct.K = s->B;
    ct.start[0] = 0;
    for (int l = 0; l < L; ++l) {
        ct.g[l] = s->DT[l];
        ct.out[l] = s->g + s->off[l] + (int64_t)s->out[l] * s->in[l];
        ct.boff[l] = s->off[l] + (int64_t)s->out[l] * s->in[l];
        ct.rows[l] = s->out[l];
        ct.start[l + 1] = ct.start[l] + s->out[l];
    }
    const dim3 ggrid((unsigned)(gt.start[L] + (ct.start[L] + 3) / 4));
    HIP_CHECK(afx::launch_ok(ggrid, dim3(256), 256));
    k_grads<<<ggrid, 256, 0, s->stream>>>(gt, ct, ad, s->sr);
    HIP_CHECK(hipGetLastError());
    return AFX_OK;
}

int adam(afx_mlp* s, float lr, bool update)
{
    AdamTable at{};
    at.n = s->L;
    for (int l = 0; l <= s->L; ++l) at.off[l] = s->off[l];
    at.tstart[0] = 0;
    at.bofs[0] = 0;
    for (int l = 0; l < s->L; ++l) {
        at.rows[l] = s->out[l];
        at.cols[l] = s->in[l];
        at.wb[l] = s->wb[l];
        at.wt[l] = s->wt[l];
        at.tstart[l + 1] = at.tstart[l] + (s->out[l] / 32) * (s->in[l] / 32);
        at.bofs[l + 1] = at.bofs[l] + s->out[l];
    }
    float bc1 = 1.0f, bc2 = 1.0f;
    if (update) {
        s->t += 1;
        bc1 = (float)(1.0 - std::pow((double)s->cfg.beta1, (double)s->t));
        bc2 = (float)(1.0 - std::pow((double)s->cfg.beta2, (double)s->t));
    }
    const int64_t blocks = at.tstart[s->L] + (at.bofs[s->L] + 255) / 256;
    HIP_CHECK(afx::launch_ok(dim3((unsigned)blocks), dim3(256), 256));
    k_adam<<<(unsigned)blocks, 256, 0, s->stream>>>(at, s->w, s->m1, s->m2, s->g, lr, s->cfg.beta1, s->cfg.beta2,
                                                    s->cfg.eps, s->cfg.weight_decay, bc1, bc2, update ? 1 : 0, s->sr);
    HIP_CHECK(hipGetLastError());
    return AFX_OK;
}

int set_device(afx_mlp* s)
{
    HIP_CHECK(hipSetDevice(s->device));
    return AFX_OK;
}

// Device address of `scratch` on each device, looked up once per device
// with that device current.
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_dev[kMaxScratchDevices] = {};

int scratch_setup(afx_mlp* s)
{
    if (s->device >= kMaxScratchDevices) return err(AFX_E_RANGE, "device %d above %d", s->device, kMaxScratchDevices - 1);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[s->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return err(AFX_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", s->device, hipGetErrorString(e));
        g_scratch_dev[s->device] = static_cast<unsigned char*>(p);
    }
    s->scratch_dev = g_scratch_dev[s->device];
    return AFX_OK;
}

// Byte offsets of one training call's buffers (each rounded up to 256):
// minibatch indices [steps][B] int32, loss partials [steps][nblk_loss],
// schedule [steps][3] FP32, step counter int64.  Returns the total.
size_t train_layout(const afx_mlp* s, int64_t steps, size_t off[4])
{
    auto up = [](size_t n) { return (n + 255) & ~size_t(255); };
    off[0] = 0;
    off[1] = off[0] + up((size_t)steps * s->B * sizeof(int));
    off[2] = off[1] + up((size_t)steps * s->nblk_loss * sizeof(float));
    off[3] = off[2] + up((size_t)steps * 3 * sizeof(float));
    return off[3] + 256;
}

// Minibatch rows (a fresh seeded permutation per epoch, consumed B at a
// time) and, per step, the cosine learning rate from lr0 to lr1 and Adam's
// bias corrections, into idx[steps*B] and sched[steps*3].
void make_schedule(afx_mlp* s, int64_t steps, float lr0, float lr1, uint32_t seed, int* idx, float* sched)
{
    std::vector<int>& perm = s->perm;
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937_64 rng(seed);
    size_t pos = perm.size();
    const size_t n = (size_t)steps * s->B;
    for (size_t i = 0; i < n; ++i) {
        if (pos == perm.size()) {
            std::shuffle(perm.begin(), perm.end(), rng);
            pos = 0;
        }
        idx[i] = perm[pos++];
    }
    const int64_t t0 = s->t;
    for (int64_t k = 0; k < steps; ++k) {
        const double frac = steps > 1 ? (double)k / (double)(steps - 1) : 1.0;
        sched[(size_t)(3 * k)] = (float)(lr1 + 0.5 * (lr0 - lr1) * (1.0 + std::cos(M_PI * frac)));
        sched[(size_t)(3 * k + 1)] = (float)(1.0 - std::pow((double)s->cfg.beta1, (double)(t0 + k + 1)));
        sched[(size_t)(3 * k + 2)] = (float)(1.0 - std::pow((double)s->cfg.beta2, (double)(t0 + k + 1)));
    }
}

// Host image of the scratch layout for afx_mlp_train_noalloc (static, so the
// call allocates nothing); guarded by g_stage_mu for the whole call.
std::mutex g_stage_mu;
alignas(256) unsigned char g_stage[AFX_MLP_TRAIN_SCRATCH_BYTES];

void init_params(const afx_mlp* s, std::vector<float>& p)
{
    std::mt19937_64 rng(s->cfg.seed);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    p.assign((size_t)s->nparam, 0.0f);
    for (int l = 0; l < s->L; ++l) {
        // He-uniform for the SiLU layers, a tenth of it for the linear output
        double a = std::sqrt(6.0 / s->in[l]);
        if (l == s->L - 1) a *= 0.1;
        float* W = p.data() + s->off[l];
        for (int64_t i = 0; i < (int64_t)s->out[l] * s->in[l]; ++i) W[i] = (float)(a * u(rng));
    }
}

}  // namespace

namespace {

constexpr int kInferRows = 32768;

int ensure_infer(afx_mlp* s)
{
    if (s->imem) return AFX_OK;
    int wmax = 32;
    for (int l = 0; l + 1 < s->L; ++l) wmax = std::max(wmax, s->out[l]);
    Arena probe;
    auto layout = [&](Arena& a) {
        s->Xi = a.take<bf16>((size_t)kInferRows * s->F);
        s->Ai[0] = a.take<bf16>((size_t)kInferRows * wmax);
        s->Ai[1] = a.take<bf16>((size_t)kInferRows * wmax);
        s->Yi = a.take<float>((size_t)kInferRows * s->cfg.out_dim);
        s->xqi = a.take<float>((size_t)kInferRows * s->cfg.in_dim);
    };
    layout(probe);
    if (hipMalloc(&s->imem, probe.used) != hipSuccess)
        return err(AFX_E_NOMEM, "device allocation of %zu bytes for inference failed", probe.used);
    Arena real;
    real.base = static_cast<char*>(s->imem);
    real.cap = probe.used;
    layout(real);
    return AFX_OK;
}

// Network outputs for the m rows (a multiple of 32) staged in xqi: features,
// hidden layers through the ping-pong buffers, output layer into y (row
// stride out_dim).  Each output row depends on its input row only, with the
// same K order as the batch-sized training kernels, so chunking does not
// change a value.
int forward_chunk(afx_mlp* s, int m, float* y)
{
    const int L = s->L;
    HIP_CHECK(launch_features(s, s->xqi, nullptr, m, s->Xi, nullptr, m, StepRef{}));
    const bf16* act = s->Xi;
    for (int l = 0; l < L; ++l) {
        Gemm g{};
        g.A = act;
        g.lda = s->in[l];
        g.B = s->wb[l];
        g.ldb = s->in[l];
        g.M = m;
        g.N = s->out[l];
        g.K = s->in[l];
        g.bias = s->w + s->off[l] + (int64_t)s->out[l] * s->in[l];
        if (l + 1 < L) {
            g.act = s->Ai[l & 1];
            g.ldact = s->out[l];
            if (int rc = gemm_launch(s, EPI_HIDDEN, g)) return rc;
            act = s->Ai[l & 1];
        } else {
            g.f32 = y;
            g.ldf = s->out[l];
            if (int rc = gemm_launch(s, EPI_OUT, g)) return rc;
        }
    }
    return AFX_OK;
}

// Both forward entry points: x and y are host (dev = false) or device memory.
int forward_rows(afx_mlp* s, const float* x, int64_t rows, float* y, bool dev)
{
    if (int rc = set_device(s)) return rc;
    if (int rc = ensure_infer(s)) return rc;
    const int I = s->cfg.in_dim, O = s->cfg.out_dim;
    const auto t0 = std::chrono::steady_clock::now();
    HIP_CHECK(hipEventRecord(s->ev0, s->stream));
    for (int64_t r0 = 0; r0 < rows; r0 += kInferRows) {
        const int64_t nr = std::min<int64_t>(kInferRows, rows - r0);
        const int m = (int)((nr + 63) / 64 * 64);
        if (nr < m) HIP_CHECK(hipMemsetAsync(s->xqi + nr * I, 0, (size_t)(m - nr) * I * sizeof(float), s->stream));
        HIP_CHECK(hipMemcpyAsync(s->xqi, x + r0 * I, (size_t)nr * I * sizeof(float),
                                 dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s->stream));
        // a full device chunk is written in place; otherwise through Yi
        const bool direct = dev && nr == m;
        if (int rc = forward_chunk(s, m, direct ? y + r0 * O : s->Yi)) return rc;
        if (!direct)
            HIP_CHECK(hipMemcpyAsync(y + r0 * O, s->Yi, (size_t)nr * O * sizeof(float),
                                     dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s->stream));
    }
    HIP_CHECK(hipEventRecord(s->ev1, s->stream));
    HIP_CHECK(hipStreamSynchronize(s->stream));
    float ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->st.last_forward_gpu_ms = ms;
    s->st.last_forward_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    s->st.last_forward_rows = rows;
    return AFX_OK;
}

}  // namespace

extern "C" {

int afx_mlp_config_default(afx_mlp_config* c)
{
    if (!c) return err(AFX_E_ARG, "null config");
    *c = afx_mlp_config{};
    c->in_dim = 6;
    c->n_freq = 64;
    c->width = 512;
    c->depth = 6;
    c->out_dim = 1024;
    c->batch = 256;
    c->device = 0;
    c->seed = 1;
    c->freq_min = 0.25f;
    c->freq_max = 32.0f;
    c->beta1 = 0.9f;
    c->beta2 = 0.999f;
    c->eps = 1e-8f;
    c->weight_decay = 0.0f;
    c->flags = 0;
    return AFX_OK;
}

int afx_mlp_create(const afx_mlp_config* cfg, afx_mlp** out)
{
    if (!cfg || !out) return err(AFX_E_ARG, "null argument");
    *out = nullptr;
    const afx_mlp_config& c = *cfg;
    if (c.in_dim < 1 || c.n_freq < 1 || c.depth < 2 || c.depth > kMaxLayers)
        return err(AFX_E_ARG, "need in_dim >= 1, n_freq >= 1 and 2 <= depth <= %d", kMaxLayers);
    if (c.in_dim * c.n_freq > 1024)
        return err(AFX_E_ARG, "in_dim * n_freq = %d above 1024 (one feature workgroup per row)", c.in_dim * c.n_freq);
    const int F = 2 * c.in_dim * c.n_freq;
    if (F % 32 || c.width % 32 || c.out_dim % 32 || c.batch % 32 || c.width < 32 || c.out_dim < 32 || c.batch < 32)
        return err(AFX_E_ARG, "features (%d), width, out_dim and batch must be positive multiples of 32", F);
    if ((int64_t)c.batch * std::max(F, std::max(c.width, c.out_dim)) > (int64_t)1 << 30)
        return err(AFX_E_RANGE, "batch x layer size above 2^30");
    if (!(c.freq_min > 0) || !(c.freq_max >= c.freq_min)) return err(AFX_E_ARG, "need 0 < freq_min <= freq_max");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return err(AFX_E_HIP, "no HIP device available");
    if (c.device < 0 || c.device >= n) return err(AFX_E_ARG, "device %d outside 0..%d", c.device, n - 1);

    afx_mlp* s = new (std::nothrow) afx_mlp;
    if (!s) return err(AFX_E_NOMEM, "out of host memory");
    s->cfg = c;
    s->device = c.device;
    s->L = c.depth;
    s->F = F;
    s->B = c.batch;
    for (int l = 0; l < s->L; ++l) {
        s->in.push_back(l == 0 ? F : c.width);
        s->out.push_back(l == s->L - 1 ? c.out_dim : c.width);
    }
    s->off.assign(s->L + 1, 0);
    for (int l = 0; l < s->L; ++l) s->off[l + 1] = s->off[l] + (int64_t)s->out[l] * s->in[l] + s->out[l];
    s->nparam = s->off[s->L];
    s->freq_host.resize(c.n_freq);
    for (int k = 0; k < c.n_freq; ++k) {
        const double r = c.n_freq > 1 ? (double)k / (c.n_freq - 1) : 0.0;
        s->freq_host[k] = (float)(c.freq_min * std::pow((double)c.freq_max / c.freq_min, r));
    }

    auto fail_free = [&](int rc) {
        afx_mlp_destroy(s);
        return rc;
    };
    if (int rc = set_device(s)) return fail_free(rc);
    if (hipStreamCreate(&s->stream) != hipSuccess || hipEventCreate(&s->ev0) != hipSuccess ||
        hipEventCreate(&s->ev1) != hipSuccess)
        return fail_free(err(AFX_E_HIP, "stream/event creation failed"));
    if (int rc = scratch_setup(s)) return fail_free(rc);

    const int B = s->B;
    s->nblk_loss = (B / 32) * (c.out_dim / 32);
    // sizes of everything, then one allocation
    Arena probe;
    auto layout = [&](Arena& a) {
        s->w = a.take<float>(s->nparam);
        s->m1 = a.take<float>(s->nparam);
        s->m2 = a.take<float>(s->nparam);
        s->g = a.take<float>(s->nparam);
        s->freq = a.take<float>(c.n_freq);
        s->wb.assign(s->L, nullptr);
        s->wt.assign(s->L, nullptr);
        for (int l = 0; l < s->L; ++l) {
            s->wb[l] = a.take<bf16>((size_t)s->out[l] * s->in[l]);
            s->wt[l] = a.take<bf16>((size_t)s->out[l] * s->in[l]);
        }
        s->X = a.take<bf16>((size_t)B * F);
        s->XT = a.take<bf16>((size_t)B * F);
        s->Z.assign(s->L, nullptr);
        s->A.assign(s->L, nullptr);
        s->AT.assign(s->L, nullptr);
        s->D.assign(s->L, nullptr);
        s->DT.assign(s->L, nullptr);
        for (int l = 0; l < s->L; ++l) {
            if (l + 1 < s->L) {
                s->Z[l] = a.take<float>((size_t)B * s->out[l]);
                s->A[l] = a.take<bf16>((size_t)B * s->out[l]);
                s->AT[l] = a.take<bf16>((size_t)B * s->out[l]);
            }
            s->D[l] = a.take<bf16>((size_t)B * s->out[l]);
            s->DT[l] = a.take<bf16>((size_t)B * s->out[l]);
        }
        s->Y = a.take<float>((size_t)B * c.out_dim);
        s->xq = a.take<float>((size_t)B * c.in_dim);
        s->tq = a.take<float>((size_t)B * c.out_dim);
        s->wq = a.take<float>(B);
        s->partial = a.take<float>(s->nblk_loss);
    };
    layout(probe);
    if (hipMalloc(&s->mem, probe.used) != hipSuccess) return fail_free(err(AFX_E_NOMEM, "device allocation of %zu bytes failed", probe.used));
    Arena real;
    real.base = static_cast<char*>(s->mem);
    real.cap = probe.used;
    layout(real);
    if (hipMemsetAsync(s->mem, 0, probe.used, s->stream) != hipSuccess)
        return fail_free(err(AFX_E_HIP, "clearing device memory failed"));
    if (hipMemcpyAsync(s->freq, s->freq_host.data(), c.n_freq * sizeof(float), hipMemcpyHostToDevice, s->stream) !=
        hipSuccess)
        return fail_free(err(AFX_E_HIP, "frequency upload failed"));
    std::vector<float> p;
    init_params(s, p);
    if (int rc = afx_mlp_set_params(s, p.data(), (int64_t)p.size())) return fail_free(rc);
    s->st.params = s->nparam;
    double fwd = 0, bwd = 0;
    for (int l = 0; l < s->L; ++l) {
        fwd += 2.0 * B * s->in[l] * s->out[l];
        bwd += 2.0 * B * s->in[l] * s->out[l] * (l == 0 ? 1 : 2);
    }
    s->st.flop_per_step = fwd + bwd;
    *out = s;
    return AFX_OK;
}

void afx_mlp_destroy(afx_mlp* s)
{
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->mem) (void)hipFree(s->mem);
    if (s->imem) (void)hipFree(s->imem);
    if (s->xd) (void)hipFree(s->xd);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int afx_mlp_param_count(afx_mlp* s, int64_t* n)
{
    if (!s || !n) return err(AFX_E_ARG, "null argument");
    *n = s->nparam;
    return AFX_OK;
}

int afx_mlp_get_params(afx_mlp* s, float* out, int64_t cap)
{
    if (!s || !out) return err(AFX_E_ARG, "null argument");
    if (cap < s->nparam) return err(AFX_E_RANGE, "buffer of %lld floats, %lld needed", (long long)cap, (long long)s->nparam);
    if (int rc = set_device(s)) return rc;
    HIP_CHECK(hipMemcpyAsync(out, s->w, s->nparam * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIP_CHECK(hipStreamSynchronize(s->stream));
    return AFX_OK;
}

int afx_mlp_set_params(afx_mlp* s, const float* in, int64_t n)
{
    if (!s || !in) return err(AFX_E_ARG, "null argument");
    if (n != s->nparam) return err(AFX_E_ARG, "%lld parameters given, %lld expected", (long long)n, (long long)s->nparam);
    if (int rc = set_device(s)) return rc;
    HIP_CHECK(hipMemcpyAsync(s->w, in, n * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_CHECK(hipMemsetAsync(s->m1, 0, n * sizeof(float), s->stream));
    HIP_CHECK(hipMemsetAsync(s->m2, 0, n * sizeof(float), s->stream));
    s->t = 0;
    if (int rc = adam(s, 0.0f, false)) return rc;
    HIP_CHECK(hipStreamSynchronize(s->stream));
    return AFX_OK;
}

int afx_mlp_get_frequencies(afx_mlp* s, float* out, int32_t cap)
{
    if (!s || !out) return err(AFX_E_ARG, "null argument");
    if (cap < s->cfg.n_freq) return err(AFX_E_RANGE, "buffer of %d floats, %d needed", cap, s->cfg.n_freq);
    std::memcpy(out, s->freq_host.data(), s->freq_host.size() * sizeof(float));
    return AFX_OK;
}

int afx_mlp_forward(afx_mlp* s, const float* x, int64_t rows, float* y)
{
    if (!s || !x || !y || rows < 1) return err(AFX_E_ARG, "forward needs x, y and rows >= 1");
    return forward_rows(s, x, rows, y, false);
}

int afx_mlp_forward_device(afx_mlp* s, const float* x, int64_t rows, float* y)
{
    if (!s || !x || !y || rows < 1) return err(AFX_E_ARG, "forward needs x, y and rows >= 1");
    return forward_rows(s, x, rows, y, true);
}

int afx_mlp_gradients(afx_mlp* s, const float* x, const float* target, const float* weight, int64_t rows,
                      float* grad, double* loss, float* y)
{
    if (!s || !x || !target || !weight || !grad) return err(AFX_E_ARG, "null argument");
    if (rows < 1 || rows > s->B) return err(AFX_E_RANGE, "rows %lld outside 1..%d", (long long)rows, s->B);
    if (int rc = set_device(s)) return rc;
    const int B = s->B, I = s->cfg.in_dim, O = s->cfg.out_dim;
    // padding rows have zero position, target and weight, so they add nothing to loss or gradient
    HIP_CHECK(hipMemsetAsync(s->xq, 0, (size_t)B * I * sizeof(float), s->stream));
    HIP_CHECK(hipMemsetAsync(s->tq, 0, (size_t)B * O * sizeof(float), s->stream));
    HIP_CHECK(hipMemsetAsync(s->wq, 0, (size_t)B * sizeof(float), s->stream));
    HIP_CHECK(hipMemcpyAsync(s->xq, x, (size_t)rows * I * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_CHECK(hipMemcpyAsync(s->tq, target, (size_t)rows * O * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_CHECK(hipMemcpyAsync(s->wq, weight, (size_t)rows * sizeof(float), hipMemcpyHostToDevice, s->stream));
    if (int rc = step_grad(s, s->xq, s->tq, s->wq, nullptr, s->Y)) return rc;
    std::vector<float> part((size_t)s->nblk_loss);
    HIP_CHECK(hipMemcpyAsync(grad, s->g, s->nparam * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIP_CHECK(hipMemcpyAsync(part.data(), s->partial, part.size() * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    if (y) HIP_CHECK(hipMemcpyAsync(y, s->Y, (size_t)rows * O * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIP_CHECK(hipStreamSynchronize(s->stream));
    if (loss) {
        double acc = 0.0;
        for (float v : part) acc += v;
        *loss = acc;
    }
    return AFX_OK;
}

int afx_mlp_set_data(afx_mlp* s, const float* x, const float* target, const float* weight, int64_t rows)
{
    if (!s || !x || !target || !weight) return err(AFX_E_ARG, "null argument");
    if (rows < s->B) return err(AFX_E_ARG, "training set of %lld rows is smaller than one batch (%d)", (long long)rows, s->B);
    if (rows > INT32_MAX / 2) return err(AFX_E_RANGE, "training set too large");
    if (int rc = set_device(s)) return rc;
    HIP_CHECK(hipStreamSynchronize(s->stream));
    if (s->xd) HIP_CHECK(hipFree(s->xd));
    s->xd = s->td = s->wd = nullptr;
    s->ndata = 0;
    const int I = s->cfg.in_dim, O = s->cfg.out_dim;
    Arena probe;
    auto layout = [&](Arena& a) {
        float* xd = a.take<float>((size_t)rows * I);
        float* td = a.take<float>((size_t)rows * O);
        float* wd = a.take<float>((size_t)rows);
        return std::make_tuple(xd, td, wd);
    };
    layout(probe);
    void* mem = nullptr;
    if (hipMalloc(&mem, probe.used) != hipSuccess) return err(AFX_E_NOMEM, "device allocation of %zu bytes failed", probe.used);
    Arena real;
    real.base = static_cast<char*>(mem);
    std::tie(s->xd, s->td, s->wd) = layout(real);
    HIP_CHECK(hipMemcpyAsync(s->xd, x, (size_t)rows * I * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_CHECK(hipMemcpyAsync(s->td, target, (size_t)rows * O * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_CHECK(hipMemcpyAsync(s->wd, weight, (size_t)rows * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_CHECK(hipStreamSynchronize(s->stream));
    s->perm.assign((size_t)rows, 0);
    s->ndata = rows;
    return AFX_OK;
}

int afx_mlp_train(afx_mlp* s, int64_t steps, float lr0, float lr1, uint32_t seed, float* loss_out)
{
    if (!s) return err(AFX_E_ARG, "null handle");
    if (steps < 1) return err(AFX_E_ARG, "steps must be >= 1");
    if (!s->xd) return err(AFX_E_STATE, "no training set (afx_mlp_set_data)");
    if (!(lr0 >= 0) || !(lr1 >= 0)) return err(AFX_E_ARG, "learning rates must be >= 0");
    if (int rc = set_device(s)) return rc;
    const int B = s->B;
    std::vector<int> idx((size_t)steps * B);
    std::vector<float> sched((size_t)steps * 3);
    make_schedule(s, steps, lr0, lr1, seed, idx.data(), sched.data());
    const int64_t t0 = s->t;
    int* d_idx = nullptr;
    float *d_part = nullptr, *d_sched = nullptr;
    int64_t* d_step = nullptr;
    hipGraphExec_t exec = nullptr;
    float* keep_partial = s->partial;
    auto cleanup = [&](int rc) {
        (void)hipStreamSynchronize(s->stream);
        if (exec) (void)hipGraphExecDestroy(exec);
        (void)hipFree(d_idx);
        (void)hipFree(d_part);
        (void)hipFree(d_sched);
        (void)hipFree(d_step);
        s->partial = keep_partial;
        s->sr = StepRef{};
        return rc;
    };
    if (hipMalloc(&d_idx, idx.size() * sizeof(int)) != hipSuccess ||
        hipMalloc(&d_part, (size_t)steps * s->nblk_loss * sizeof(float)) != hipSuccess ||
        hipMalloc(&d_sched, sched.size() * sizeof(float)) != hipSuccess || hipMalloc(&d_step, sizeof(int64_t)) != hipSuccess)
        return cleanup(err(AFX_E_NOMEM, "device allocation for the training schedule failed"));
    if (hipMemcpyAsync(d_idx, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice, s->stream) != hipSuccess ||
        hipMemcpyAsync(d_sched, sched.data(), sched.size() * sizeof(float), hipMemcpyHostToDevice, s->stream) !=
            hipSuccess ||
        hipMemsetAsync(d_step, 0, sizeof(int64_t), s->stream) != hipSuccess)
        return cleanup(err(AFX_E_HIP, "training schedule upload failed"));
    s->partial = d_part;
    s->sr = StepRef{d_step, (int64_t)B, (int64_t)s->nblk_loss, d_sched};
    // One step (forward, loss, backward, AdamW) at step counter + rel; a
    // captured chunk advances the counter once, after its last step.
    auto enqueue_step = [&](int64_t rel) -> int {
        s->sr.rel = rel;
        return step_grad(s, s->xd, s->td, s->wd, d_idx, nullptr, true);
    };
    // Steps are replayed from a captured graph of kChunk steps; the rest run eagerly.
    constexpr int64_t kChunk = 32;
    const int64_t nfull = (s->cfg.flags & AFX_MLP_NO_GRAPH) ? 0 : steps / kChunk;
    if (nfull > 0) {
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return cleanup(err(AFX_E_HIP, "graph capture failed to start"));
        int rc = AFX_OK;
        for (int64_t k = 0; k < kChunk && !rc; ++k) rc = enqueue_step(k);
        if (!rc) {
            k_step_inc<<<1, 64, 0, s->stream>>>(d_step, kChunk);
            if (hipGetLastError() != hipSuccess) rc = err(AFX_E_HIP, "step counter launch failed");
        }
        const hipError_t ce = hipStreamEndCapture(s->stream, &g);
        if (rc || ce != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return cleanup(rc ? rc : err(AFX_E_HIP, "graph capture failed"));
        }
        const hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) return cleanup(err(AFX_E_HIP, "graph instantiation failed"));
    }
    const auto w0 = std::chrono::steady_clock::now();
    if (hipEventRecord(s->ev0, s->stream) != hipSuccess) return cleanup(err(AFX_E_HIP, "event record failed"));
    for (int64_t c = 0; c < nfull; ++c)
        if (hipGraphLaunch(exec, s->stream) != hipSuccess) return cleanup(err(AFX_E_HIP, "graph launch failed"));
    // the rest eagerly, at the counter the chunks left (nfull * kChunk)
    for (int64_t k = nfull * kChunk; k < steps; ++k)
        if (int rc = enqueue_step(k - nfull * kChunk)) return cleanup(rc);
    if (hipEventRecord(s->ev1, s->stream) != hipSuccess) return cleanup(err(AFX_E_HIP, "event record failed"));
    if (hipStreamSynchronize(s->stream) != hipSuccess) return cleanup(err(AFX_E_HIP, "training failed on the device"));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, s->ev0, s->ev1);
    s->t = t0 + steps;
    s->st.last_train_gpu_ms = ms;
    s->st.last_train_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    s->st.train_steps = s->t;
    if (loss_out) {
        std::vector<float> part((size_t)steps * s->nblk_loss);
        if (hipMemcpy(part.data(), d_part, part.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            return cleanup(err(AFX_E_HIP, "loss readback failed"));
        for (int64_t k = 0; k < steps; ++k) {
            double acc = 0.0;
            for (int b = 0; b < s->nblk_loss; ++b) acc += part[(size_t)k * s->nblk_loss + b];
            loss_out[k] = (float)acc;
        }
    }
    return cleanup(AFX_OK);
}

int afx_mlp_train_noalloc(afx_mlp* s, int64_t steps, float lr0, float lr1, uint32_t seed, float* loss_out)
{
    if (!s) return err(AFX_E_ARG, "null handle");
    if (steps < 1) return err(AFX_E_ARG, "steps must be >= 1");
    if (!s->xd) return err(AFX_E_STATE, "no training set (afx_mlp_set_data)");
    if (!(lr0 >= 0) || !(lr1 >= 0)) return err(AFX_E_ARG, "learning rates must be >= 0");
    if (!s->scratch_dev) return err(AFX_E_STATE, "training scratch not set up on device %d", s->device);
    size_t off[4];
    if (steps > (int64_t)AFX_MLP_TRAIN_SCRATCH_BYTES)
        return err(AFX_E_RANGE, "%lld steps exceed the %u-byte training scratch", (long long)steps,
                   (unsigned)AFX_MLP_TRAIN_SCRATCH_BYTES);
    const size_t need = train_layout(s, steps, off);
    if (need > AFX_MLP_TRAIN_SCRATCH_BYTES)
        return err(AFX_E_RANGE,
                   "%lld steps at batch %d need %zu bytes of training scratch, above its %u bytes: split the call",
                   (long long)steps, s->B, need, (unsigned)AFX_MLP_TRAIN_SCRATCH_BYTES);
    if (int rc = set_device(s)) return rc;
    std::lock_guard<std::mutex> lk(g_stage_mu);
    make_schedule(s, steps, lr0, lr1, seed, reinterpret_cast<int*>(g_stage + off[0]),
                  reinterpret_cast<float*>(g_stage + off[2]));
    unsigned char* base = s->scratch_dev;
    int* d_idx = reinterpret_cast<int*>(base + off[0]);
    float* d_part = reinterpret_cast<float*>(base + off[1]);
    float* d_sched = reinterpret_cast<float*>(base + off[2]);
    int64_t* d_step = reinterpret_cast<int64_t*>(base + off[3]);
    const int64_t t0 = s->t;
    float* keep_partial = s->partial;
    auto finish = [&](int rc) {
        (void)hipStreamSynchronize(s->stream);
        s->partial = keep_partial;
        s->sr = StepRef{};
        return rc;
    };
    if (hipMemcpyAsync(d_idx, g_stage + off[0], off[1] - off[0], hipMemcpyHostToDevice, s->stream) != hipSuccess ||
        hipMemcpyAsync(d_sched, g_stage + off[2], off[3] - off[2], hipMemcpyHostToDevice, s->stream) != hipSuccess)
        return finish(err(AFX_E_HIP, "training schedule upload failed"));
    k_step_zero<<<1, 64, 0, s->stream>>>(d_step);
    if (hipGetLastError() != hipSuccess) return finish(err(AFX_E_HIP, "step counter launch failed"));
    s->partial = d_part;
    s->sr = StepRef{d_step, (int64_t)s->B, (int64_t)s->nblk_loss, d_sched};
    const auto w0 = std::chrono::steady_clock::now();
    if (hipEventRecord(s->ev0, s->stream) != hipSuccess) return finish(err(AFX_E_HIP, "event record failed"));
    for (int64_t k = 0; k < steps; ++k) {
        s->sr.rel = k;
        if (int rc = step_grad(s, s->xd, s->td, s->wd, d_idx, nullptr, true)) return finish(rc);
    }
    if (hipEventRecord(s->ev1, s->stream) != hipSuccess) return finish(err(AFX_E_HIP, "event record failed"));
    if (hipStreamSynchronize(s->stream) != hipSuccess) return finish(err(AFX_E_HIP, "training failed on the device"));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, s->ev0, s->ev1);
    s->t = t0 + steps;
    s->st.last_train_gpu_ms = ms;
    s->st.last_train_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    s->st.train_steps = s->t;
    if (loss_out) {
        float* part = reinterpret_cast<float*>(g_stage + off[1]);
        if (hipMemcpy(part, d_part, (size_t)steps * s->nblk_loss * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            return finish(err(AFX_E_HIP, "loss readback failed"));
        for (int64_t k = 0; k < steps; ++k) {
            double acc = 0.0;
            for (int b = 0; b < s->nblk_loss; ++b) acc += part[(size_t)k * s->nblk_loss + b];
            loss_out[k] = (float)acc;
        }
    }
    return finish(AFX_OK);
}

int afx_mlp_get_stats(afx_mlp* s, afx_mlp_stats* st)
{
    if (!s || !st) return err(AFX_E_ARG, "null argument");
    *st = s->st;
    st->train_steps = s->t;
    return AFX_OK;
}

}  // extern "C"
