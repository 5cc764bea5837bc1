/*
 * pmmlp.h -- kernels of the spectral surrogate: a small MLP (dense layers,
 * GELU) trained and evaluated on the matrix cores in BF16 with FP32
 * accumulation and FP32 master weights.
 *
 * The reference has no learned model; its design loop is the parameter
 * sweep of src/prismo/optimization/sweep.py:119-167 (every candidate a full
 * FDTD run).  These entry points let a caller screen candidates with a
 * network trained on spectra from the batched FDTD (pmbatch.h) and send
 * only the promising or uncertain ones back to FDTD.
 *
 * All pointers are device memory (FP32, row-major with the given strides),
 * work is enqueued on `stream` (hipStream_t); status codes as pmfdtd.h.
 */
#ifndef PMMLP_H
#define PMMLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum pm_mlp_mode {
  PM_MLP_LINEAR = 0,  /* C = A B (+ bias[n])                                 */
  PM_MLP_GELU = 1,    /* Z = A B + bias[n] stored to Z;  C = gelu(Z)         */
  PM_MLP_DGELU = 2,   /* C = (A B) * gelu'(Z)   (Z read: the layer's input)  */
  PM_MLP_PLAIN = 3    /* C = A B                                            */
};

/* C[m][n] (m < M, n < N, row stride ldc) from A[m][k] = A[m*lda_m + k*lda_k]
 * and B[k][n] = B[k*ldb_k + n*ldb_n], k < K.  Operands are rounded to BF16
 * (round to nearest even) for the MFMA products, sums are FP32.
 * gelu(z) = z (1 + erf(z / sqrt 2)) / 2. */
int pm_mlp_gemm(void* stream, int mode, int32_t M, int32_t N, int32_t K, const float* A,
                int64_t lda_m, int64_t lda_k, const float* B, int64_t ldb_k, int64_t ldb_n,
                float* C, int64_t ldc, const float* bias, float* Z);

/* out[n] = sum_m A[m*lda + n], n < N (bias gradients). */
int pm_mlp_colsum(void* stream, int32_t M, int32_t N, const float* A, int64_t lda, float* out);

/* Mean squared error over M x N with a per-column weight w[n] (NULL = 1):
 * loss[0] = sum w (Y-T)^2 / (M N) (accumulated with atomics into loss, which
 * the caller zeroes), dY = 2 w (Y - T) / (M N). */
int pm_mlp_mse(void* stream, int32_t M, int32_t N, const float* Y, const float* T,
               const float* w, float* dY, float* loss);

/* Batched forms: `batch` independent problems of the same shape (1 <=
 * batch <= 65535), problem q reading and writing its operands at q times
 * the given stride (in elements) from the base pointers; each problem's
 * result is the one the single call gives.  Loss of problem q in loss[q].
 * No counterpart in the reference: the surrogate trains its ensemble
 * members in lockstep, one launch for all members. */
int pm_mlp_gemm_batched(void* stream, int mode, int32_t batch, int32_t M, int32_t N, int32_t K,
                        const float* A, int64_t lda_m, int64_t lda_k, int64_t stride_a,
                        const float* B, int64_t ldb_k, int64_t ldb_n, int64_t stride_b,
                        float* C, int64_t ldc, int64_t stride_c, const float* bias,
                        int64_t stride_bias, float* Z, int64_t stride_z);
int pm_mlp_colsum_batched(void* stream, int32_t batch, int32_t M, int32_t N, const float* A,
                          int64_t lda, int64_t stride_a, float* out, int64_t stride_out);
int pm_mlp_mse_batched(void* stream, int32_t batch, int32_t M, int32_t N, const float* Y,
                       const float* T, int64_t stride, const float* w, float* dY, float* loss);

/* Adam with decoupled weight decay on n FP32 parameters, step t >= 1. */
int pm_mlp_adam(void* stream, int64_t n, float* p, const float* g, float* m, float* v,
                float lr, float beta1, float beta2, float eps, float weight_decay, int32_t t);

#ifdef __cplusplus
}
#endif
#endif /* PMMLP_H */
