/* Learned ground-state start: a Fourier neural operator (FNO) on the
 * 128x16x16 lattice, trained and evaluated by the library's own BF16 MFMA
 * kernels (gfx950).  The reference has no learned component; this surface
 * supplies initial guesses to bec_ground_states_two_component (bec_api.h),
 * whose Thomas-Fermi start remains the fallback.
 *
 * Every operation of the network is a strided batched GEMM on BF16 operands
 * with FP32 accumulation (bec_gemm), plus element-wise kernels:
 *   lift      h0 = x P + p                       (per cell, c_in -> width)
 *   layer l   h' = act(h W_l + b_l + K_l h)      (l = 0..layers-1, no act on the last)
 *   K_l h     = Re IDFT_z IDFT_xy [ R_l(k) . DFT_xy DFT_z h ](modes)
 *             (kz in {0..mz/2-1, -mz/2..-1}, all ky, kx when my = ny, mx = nx;
 *              R_l(k) a complex width x width matrix per retained mode;
 *              both transforms scaled by 1/sqrt(cells) each)
 *   project   y = act(h Q1 + q1) Q2 + q2          (width -> hidden -> c_out)
 * act is GELU, 0.5 x (1 + erf(x / sqrt 2)).  Activations are FP32 in memory
 * and rounded to BF16 (nearest even) as GEMM operands; weights are kept as
 * FP32 masters with BF16 copies (Adam updates the masters). */
#ifndef BEC_FNO_H
#define BEC_FNO_H

#include <stddef.h>
#include <stdint.h>

#include "bec_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ GEMM
 * C[b][m][n] = alpha sum_k A[b][m][k] B[b][k][n] + beta C[b][m][n] + bias[n]
 * Index maps are two-level: off(i) = (i >> lg) * hi + (i & (2^lg - 1)) * lo,
 * in elements; a one-level map uses lg = 40.  A and B are FP32 (flag 0) or
 * BF16 (flag 1) in device memory; C is FP32.  With ksplit > 1 the K range is
 * split over ksplit workgroup layers whose partial sums go to `work`
 * (ksplit * batch * M * N floats) and are then reduced in a fixed order, so
 * results do not depend on scheduling.  M, N, K, batch >= 1. */
typedef struct bec_map2 {
    long long hi, lo;
    int lg, pad_;
} bec_map2;

typedef struct bec_gemm_desc {
    const void *A, *B;
    float *C;
    const float *bias;   /* may be NULL */
    float *work;         /* ksplit > 1 only */
    int a_bf16, b_bf16;
    int M, N, K, batch, ksplit;
    float alpha, beta;
    bec_map2 am, ak, bk, bn, cm, cn;   /* element index maps */
    bec_map2 ab, bb, cb;               /* batch index maps */
} bec_gemm_desc;

int bec_gemm(const bec_gemm_desc *d, void *hip_stream);

/* Event-timed bec_gemm on the null stream: one untimed launch, then iters
 * launches; *ms_per_gemm is their mean duration.  The operands are the
 * caller's (C is overwritten iters + 1 times). */
int bec_gemm_time(const bec_gemm_desc *d, int iters, double *ms_per_gemm);

/* Waits for all device work of the calling process (the network's kernels
 * run on the null stream unless a stream is given). */
int bec_device_synchronize(void);

/* ------------------------------------------------------------------ FNO */
typedef struct bec_fno_config {
    int nz, ny, nx;       /* lattice (cells = nz ny nx) */
    int mz, my, mx;       /* retained modes per axis (mz even, my <= ny, mx <= nx) */
    int width, layers;    /* channels, spectral layers */
    int c_in, c_out;      /* input / output channels per cell */
    int hidden;           /* projection hidden width */
    int max_batch;        /* largest batch passed to forward / train_step */
} bec_fno_config;

typedef struct bec_fno bec_fno;

int bec_fno_create(const bec_fno_config *cfg, uint64_t seed, bec_fno **out);
int bec_fno_destroy(bec_fno *f);
/* Number of FP32 parameters and the order of bec_fno_get/set_params:
 * P[c_in][width], p[width]; per layer W[width][width], b[width],
 * R[mz][my][mx][width_in][width_out][2 (re, im)]; Q1[width][hidden], q1[hidden],
 * Q2[hidden][c_out], q2[c_out]. */
long long bec_fno_param_count(const bec_fno *f);
int bec_fno_get_params(const bec_fno *f, float *host_out);
int bec_fno_set_params(bec_fno *f, const float *host_in);
/* Device buffers in the layout [batch][cells][channels] (FP32). */
int bec_fno_forward(bec_fno *f, int batch, const float *d_x, float *d_y, void *hip_stream);
/* Input features of the learned two-component ground-state start for one
 * parameter set, written on the device: d_x[cells][16] (FP32, 16-byte aligned,
 * p->cells * 16 values; cell = (z nvy + y) nvx + x).  Channels: the
 * Thomas-Fermi amplitude sqrt(max(mu - V, 0) / mu), V / mu, the coordinates
 * over their half-extents (x, y, z), the four parameters scaled into [-1, 1]
 * (`scaled`, given by the caller), and seven zeros.  The values are those of
 * the host construction in bec_amd/surrogate.py (FP64 arithmetic in the same
 * order, rounded to FP32).  No counterpart in the reference. */
int bec_fno_gs2_features(const bec_calc_params *p, const float scaled[4], float *d_x, void *hip_stream);
/* Loss = mean over batch, cells and c_out of (y - target)^2.  Fills the
 * parameter gradients (bec_fno_get_grads) and returns the loss (FP64 sum). */
int bec_fno_loss_backward(bec_fno *f, int batch, const float *d_x, const float *d_target, double *loss,
                          void *hip_stream);
int bec_fno_get_grads(const bec_fno *f, float *host_out);
/* One Adam step (beta1, beta2, eps, weight decay decoupled) on the gradients
 * of the last bec_fno_loss_backward. */
int bec_fno_adam_step(bec_fno *f, float lr, float beta1, float beta2, float eps, float weight_decay,
                      void *hip_stream);

#ifdef __cplusplus
}
#endif

#endif
