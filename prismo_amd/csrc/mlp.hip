// Surrogate MLP kernels (include/pmmlp.h): a BF16 MFMA GEMM with fused
// bias / GELU / GELU-derivative epilogues, the bias-gradient column sum,
// the weighted MSE loss and gradient, and Adam.
//
// GEMM tiling: a workgroup of 4 waves owns a 64 x 64 block of C; each wave a
// 32 x 32 quarter as 2 x 2 V_MFMA_F32_16X16X32_BF16 tiles.  K goes in blocks
// of 128: the workgroup stages A (64 x 128) and B (128 x 64) in LDS as BF16,
// both k-contiguous per row/column, so each lane reads its 8-element operand
// with one 16-byte LDS load, and runs 4 MFMA K-steps of 32 on the block.
// The next block's global loads are issued into registers before those
// MFMAs, so a GEMM of K <= 256 (every surrogate layer) waits for memory
// about once instead of once per K-step.  The global loads walk the
// unit-stride side of each operand, whichever it is (forward: X rows and W
// rows; input gradient: W columns; weight gradient: dY and X columns over
// the batch).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "../../include/pmmlp.h"
#include "../../include/pmfdtd.h"

namespace pm {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kBM = 64, kBN = 64, kBK = 128, kPad = 8;
constexpr int kStage = kBM * kBK / 256;  // elements of A (and of B) per thread per block

// FP32 -> BF16, round to nearest even (NaN kept quiet)
__device__ __forceinline__ unsigned short to_bf16(float x) {
  unsigned u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float gelu(float z) {
  return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}

// d gelu / dz = Phi(z) + z phi(z)
__device__ __forceinline__ float dgelu(float z) {
  const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
  return cdf + z * pdf;
}

// Operand strides per problem of a batch (blockIdx.z), 0 for a single GEMM.
struct GemmBatch {
  int64_t a, b, c, bias, z;
};

__global__ __launch_bounds__(256) void k_gemm(int mode, int M, int N, int K,
                                              const float* __restrict__ A, int64_t lam,
                                              int64_t lak, const float* __restrict__ B,
                                              int64_t lbk, int64_t lbn, float* __restrict__ C,
                                              int64_t ldc, const float* __restrict__ bias,
                                              float* __restrict__ Z, GemmBatch bs) {
  {
    const int64_t q = blockIdx.z;
    A += q * bs.a;
    B += q * bs.b;
    C += q * bs.c;
    if (bias) bias += q * bs.bias;
    if (Z) Z += q * bs.z;
  }
  __shared__ unsigned short sa[kBM][kBK + kPad];
  __shared__ unsigned short sb[kBN][kBK + kPad];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
  const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
  const int li = lane & 15, lq = lane >> 4;
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0, 0, 0, 0};

  const bool a_kfast = lak == 1;
  const bool b_kfast = lbk == 1;
  // element e of this thread's share of a block: (row, k) of A and (n, k) of
  // B, along the unit-stride axis of each
  auto a_rc = [&](int e, int& r, int& c) {
    const int f = tid + 256 * e;
    if (a_kfast) {
      r = f / kBK;
      c = f % kBK;
    } else {
      r = f % kBM;
      c = f / kBM;
    }
  };
  auto b_rc = [&](int e, int& r, int& c) {
    const int f = tid + 256 * e;
    if (b_kfast) {
      r = f / kBK;
      c = f % kBK;
    } else {
      r = f % kBN;
      c = f / kBN;
    }
  };
  float ra[kStage], rb[kStage];
  auto gload = [&](int k0) {
#pragma unroll
    for (int e = 0; e < kStage; ++e) {
      int r, c;
      a_rc(e, r, c);
      const int m = m0 + r, k = k0 + c;
      ra[e] = (m < M && k < K) ? A[(int64_t)m * lam + (int64_t)k * lak] : 0.0f;
    }
#pragma unroll
    for (int e = 0; e < kStage; ++e) {
      int r, c;
      b_rc(e, r, c);
      const int n = n0 + r, k = k0 + c;
      rb[e] = (n < N && k < K) ? B[(int64_t)k * lbk + (int64_t)n * lbn] : 0.0f;
    }
  };
  gload(0);
  for (int k0 = 0; k0 < K; k0 += kBK) {
#pragma unroll
    for (int e = 0; e < kStage; ++e) {
      int r, c;
      a_rc(e, r, c);
      sa[r][c] = to_bf16(ra[e]);
      b_rc(e, r, c);
      sb[r][c] = to_bf16(rb[e]);
    }
    __syncthreads();
    if (k0 + kBK < K) gload(k0 + kBK);  // in flight during this block's MFMAs
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 32) {
      bf16x8 av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const unsigned short* p = &sa[wm + 16 * a + li][kk + 8 * lq];
        av[a] = *(const bf16x8*)p;
      }
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const unsigned short* p = &sb[wn + 16 * b + li][kk + 8 * lq];
        bv[b] = *(const bf16x8*)p;
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
  // epilogue: lane holds rows 4 lq + r, column li of each 16 x 16 tile
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * a + 4 * lq + r;
        const int n = n0 + wn + 16 * b + li;
        if (m >= M || n >= N) continue;
        float v = acc[a][b][r];
        const int64_t o = (int64_t)m * ldc + n;
        if (mode == PM_MLP_LINEAR) {
          C[o] = bias ? v + bias[n] : v;
        } else if (mode == PM_MLP_GELU) {
          const float z = bias ? v + bias[n] : v;
          Z[o] = z;
          C[o] = gelu(z);
        } else if (mode == PM_MLP_DGELU) {
          C[o] = v * dgelu(Z[o]);
        } else {
          C[o] = v;
        }
      }
}

// one workgroup per 64 columns; 4 waves split the rows, then combine in LDS
__global__ __launch_bounds__(256) void k_colsum(int M, int N, const float* __restrict__ A,
                                                int64_t lda, float* __restrict__ out,
                                                int64_t sa, int64_t so) {
  __shared__ float part[4][64];
  A += (int64_t)blockIdx.y * sa;  // problem of a batch
  out += (int64_t)blockIdx.y * so;
  const int c = blockIdx.x * 64 + (threadIdx.x & 63);
  const int w = threadIdx.x >> 6;
  float s = 0.0f;
  if (c < N)
    for (int m = w; m < M; m += 4) s += A[(int64_t)m * lda + c];
  part[w][threadIdx.x & 63] = s;
  __syncthreads();
  if (w == 0 && c < N)
    out[c] = (part[0][threadIdx.x] + part[1][threadIdx.x]) +
             (part[2][threadIdx.x] + part[3][threadIdx.x]);
}

__global__ __launch_bounds__(256) void k_mse(int M, int N, const float* __restrict__ Y,
                                             const float* __restrict__ T,
                                             const float* __restrict__ w,
                                             float* __restrict__ dY, float* loss, int64_t sy) {
  __shared__ float red[256];
  {
    const int64_t q = blockIdx.y;  // problem of a batch: Y, T, dY at q sy, loss[q]
    Y += q * sy;
    T += q * sy;
    dY += q * sy;
    loss += q;
  }
  const int64_t total = (int64_t)M * N;
  const float scale = 1.0f / (float)total;
  float s = 0.0f;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    const int n = (int)(e % N);
    const float wt = w ? w[n] : 1.0f;
    const float d = Y[e] - T[e];
    s += wt * d * d;
    dY[e] = 2.0f * wt * d * scale;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0] * scale);
}

__global__ __launch_bounds__(256) void k_adam(int64_t n, float* __restrict__ p,
                                              const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v,
                                              float lr, float b1, float b2, float eps, float wd,
                                              float c1, float c2) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * 256) {
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.0f - b1) * gi;
    const float vi = b2 * v[e] + (1.0f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    const float mh = mi / c1, vh = vi / c2;
    p[e] -= lr * (mh / (sqrtf(vh) + eps) + wd * p[e]);
  }
}

int status() {
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

}  // namespace
}  // namespace pm

using namespace pm;

extern "C" {

int pm_mlp_gemm(void* stream, int mode, int32_t M, int32_t N, int32_t K, const float* A,
                int64_t lda_m, int64_t lda_k, const float* B, int64_t ldb_k, int64_t ldb_n,
                float* C, int64_t ldc, const float* bias, float* Z) {
  return pm_mlp_gemm_batched(stream, mode, 1, M, N, K, A, lda_m, lda_k, 0, B, ldb_k, ldb_n, 0,
                             C, ldc, 0, bias, 0, Z, 0);
}

int pm_mlp_gemm_batched(void* stream, int mode, int32_t batch, int32_t M, int32_t N, int32_t K,
                        const float* A, int64_t lda_m, int64_t lda_k, int64_t stride_a,
                        const float* B, int64_t ldb_k, int64_t ldb_n, int64_t stride_b,
                        float* C, int64_t ldc, int64_t stride_c, const float* bias,
                        int64_t stride_bias, float* Z, int64_t stride_z) {
  if (M < 0 || N < 0 || K < 0 || mode < 0 || mode > 3 || !A || !B || !C) return PM_EINVAL;
  if (batch < 1 || batch > 65535) return PM_EINVAL;
  if ((mode == PM_MLP_GELU || mode == PM_MLP_DGELU) && !Z) return PM_EINVAL;
  if (M == 0 || N == 0) return 0;
  dim3 grid((N + kBN - 1) / kBN, (M + kBM - 1) / kBM, batch);
  if (grid.y > 65535) return PM_EINVAL;
  const GemmBatch bs{stride_a, stride_b, stride_c, stride_bias, stride_z};
  hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, (hipStream_t)stream, mode, M, N, K, A,
                     lda_m, lda_k, B, ldb_k, ldb_n, C, ldc, bias, Z, bs);
  return status();
}

int pm_mlp_colsum(void* stream, int32_t M, int32_t N, const float* A, int64_t lda, float* out) {
  return pm_mlp_colsum_batched(stream, 1, M, N, A, lda, 0, out, 0);
}

int pm_mlp_colsum_batched(void* stream, int32_t batch, int32_t M, int32_t N, const float* A,
                          int64_t lda, int64_t stride_a, float* out, int64_t stride_out) {
  if (M < 0 || N < 0 || !A || !out || batch < 1 || batch > 65535) return PM_EINVAL;
  if (N == 0) return 0;
  hipLaunchKernelGGL(k_colsum, dim3((N + 63) / 64, batch), dim3(256), 0, (hipStream_t)stream,
                     M, N, A, lda, out, stride_a, stride_out);
  return status();
}

int pm_mlp_mse(void* stream, int32_t M, int32_t N, const float* Y, const float* T,
               const float* w, float* dY, float* loss) {
  return pm_mlp_mse_batched(stream, 1, M, N, Y, T, 0, w, dY, loss);
}

int pm_mlp_mse_batched(void* stream, int32_t batch, int32_t M, int32_t N, const float* Y,
                       const float* T, int64_t stride, const float* w, float* dY, float* loss) {
  if (M <= 0 || N <= 0 || !Y || !T || !dY || !loss || batch < 1 || batch > 65535)
    return PM_EINVAL;
  const int64_t total = (int64_t)M * N;
  const int blocks = (int)std::min<int64_t>((total + 255) / 256, 1024);
  hipLaunchKernelGGL(k_mse, dim3(blocks, batch), dim3(256), 0, (hipStream_t)stream, M, N, Y, T,
                     w, dY, loss, stride);
  return status();
}

int pm_mlp_adam(void* stream, int64_t n, float* p, const float* g, float* m, float* v,
                float lr, float beta1, float beta2, float eps, float weight_decay, int32_t t) {
  if (n < 0 || !p || !g || !m || !v || t < 1) return PM_EINVAL;
  if (n == 0) return 0;
  const float c1 = 1.0f - powf(beta1, (float)t);
  const float c2 = 1.0f - powf(beta2, (float)t);
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 2048);
  hipLaunchKernelGGL(k_adam, dim3(blocks), dim3(256), 0, (hipStream_t)stream, n, p, g, m, v,
                     lr, beta1, beta2, eps, weight_decay, c1, c2);
  return status();
}

}  // extern "C"
