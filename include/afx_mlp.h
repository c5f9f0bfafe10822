/*
 * afx surrogate — the C ABI of a learned stand-in for the FDTD impulse
 * response of one fixed room.
 *
 * No counterpart in the reference: anc-field measures every secondary path
 * by simulation (anc_field_py/ancutil.py:49-79 estimates each impulse
 * response from a full run).  This part maps a query (source position,
 * receiver position), each coordinate normalised to [0,1] by the room size,
 * to the band-limited, decimated impulse response.  The map is a
 * Fourier-feature MLP:
 *
 *   features  x -> [sin(2 pi f_k x_j), cos(2 pi f_k x_j)]  (j < in_dim, k < n_freq,
 *             geometric ladder f_k from freq_min to freq_max cycles per unit)
 *   layers    depth linear layers, width hidden units, SiLU between them,
 *             out_dim linear outputs.
 *
 * Training and inference run on hand-written BF16 MFMA kernels with FP32
 * accumulation.  Operands (features, activations, back-propagated gradients,
 * weight copies) are BF16.  Master weights, Adam moments, pre-activations and
 * outputs stay FP32.  The loss of a batch is sum_r weight_r * |y_r - t_r|^2;
 * with weight_r = 1/(rows*|t_r|^2) it is the mean squared relative L2 error.
 *
 * Parameter layout (afx_mlp_get_params / set_params / gradients): for each
 * layer l in order, W_l [out_l][in_l] row-major then b_l [out_l], FP32.
 *
 * Error handling follows afx.h: AFX_OK or a negative AFX_E_* code, with the
 * message in afx_last_error().
 */
#ifndef AFX_MLP_H
#define AFX_MLP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct afx_mlp afx_mlp;

typedef struct afx_mlp_config {
    int32_t in_dim;       /* coordinates per query (6: source xyz, receiver xyz) */
    int32_t n_freq;       /* Fourier frequencies per coordinate; 2*in_dim*n_freq must be a multiple of 32,
                             in_dim*n_freq at most 1024 */
    int32_t width;        /* hidden units per layer, multiple of 32 */
    int32_t depth;        /* linear layers, >= 2 */
    int32_t out_dim;      /* outputs per query, multiple of 32 */
    int32_t batch;        /* rows per training step and per forward chunk, multiple of 32 */
    int32_t device;
    uint32_t seed;        /* initial weights */
    float freq_min, freq_max;
    float beta1, beta2, eps, weight_decay;  /* Adam(W) */
    uint32_t flags;       /* AFX_MLP_NO_GRAPH: launch every training step eagerly */
} afx_mlp_config;

/* Training replays captured hipGraphs of 32 steps unless this flag is set. */
#define AFX_MLP_NO_GRAPH 1u

typedef struct afx_mlp_stats {
    int64_t params;
    int64_t train_steps;           /* optimizer steps taken so far */
    double flop_per_step;          /* multiply-adds x2 of one training step (forward + backward) */
    double last_train_ms;          /* wall time of the last afx_mlp_train */
    double last_train_gpu_ms;      /* device time of the same (HIP events on the handle's stream) */
    double last_forward_ms;        /* wall time of the last afx_mlp_forward */
    double last_forward_gpu_ms;
    int64_t last_forward_rows;
} afx_mlp_stats;

/* 6 -> 64 frequencies -> 6 x 512 -> 1024, batch 256, Adam(0.9, 0.999, 1e-8). */
int afx_mlp_config_default(afx_mlp_config* cfg);
int afx_mlp_create(const afx_mlp_config* cfg, afx_mlp** out);
void afx_mlp_destroy(afx_mlp* mlp);

int afx_mlp_param_count(afx_mlp* mlp, int64_t* n);
int afx_mlp_get_params(afx_mlp* mlp, float* out, int64_t cap);
/* Replaces all parameters and clears the Adam moments and step count. */
int afx_mlp_set_params(afx_mlp* mlp, const float* in, int64_t n);
/* The n_freq frequencies of the feature ladder (cycles per unit coordinate). */
int afx_mlp_get_frequencies(afx_mlp* mlp, float* out, int32_t cap);

/* y[rows][out_dim] for x[rows][in_dim]; any rows >= 1.  Rows are processed
 * in chunks of up to 32 768 with inference buffers of their own (allocated on
 * the first call); each output row depends on its input row only, so the
 * values do not depend on the chunking or on the batch size. */
int afx_mlp_forward(afx_mlp* mlp, const float* x, int64_t rows, float* y);
/* The same with x and y in device memory of the handle's device, for callers
 * whose queries and consumers live on the GPU (no host copies).  The call is
 * ordered after work on the null stream and returns when y is written. */
int afx_mlp_forward_device(afx_mlp* mlp, const float* x, int64_t rows, float* y);

/* Gradient of the weighted loss over one batch (rows <= batch), without an
 * update: grad has the parameter layout; *loss receives the loss.  When y is
 * not NULL it receives the FP32 outputs [rows][out_dim]. */
int afx_mlp_gradients(afx_mlp* mlp, const float* x, const float* target, const float* weight, int64_t rows,
                      float* grad, double* loss, float* y);

/* Training set kept in device memory: x[rows][in_dim], target[rows][out_dim], weight[rows]. */
int afx_mlp_set_data(afx_mlp* mlp, const float* x, const float* target, const float* weight, int64_t rows);

/* `steps` Adam steps on minibatches drawn without replacement from the
 * training set (a fresh seeded permutation per epoch), learning rate
 * decaying from lr0 to lr1 on a cosine over this call.  loss_out, when not
 * NULL, receives each step's batch loss (mean over the batch rows' weights as
 * given). */
int afx_mlp_train(afx_mlp* mlp, int64_t steps, float lr0, float lr1, uint32_t seed, float* loss_out);

/* afx_mlp_train without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and the same results, bit for bit, as
 * afx_mlp_train; every step is launched eagerly (no graph is built).  The
 * minibatch indices, schedule, loss partials and step counter of the call
 * live in a fixed device array of AFX_MLP_TRAIN_SCRATCH_BYTES per GPU, held
 * until the process ends, instead of buffers allocated per call.  A call
 * needs steps * (4 batch + 4 (batch/32)(out_dim/32) + 12) bytes, each of
 * the three parts rounded up to 256, plus 256; a larger call returns
 * AFX_E_RANGE (split it).  Calls on one GPU share the array (and a host
 * staging image of it): calls that overlap on one GPU race, so the library
 * serialises them within one process. */
#define AFX_MLP_TRAIN_SCRATCH_BYTES (8u << 20)
int afx_mlp_train_noalloc(afx_mlp* mlp, int64_t steps, float lr0, float lr1, uint32_t seed, float* loss_out);

int afx_mlp_get_stats(afx_mlp* mlp, afx_mlp_stats* st);

#ifdef __cplusplus
}
#endif
#endif
