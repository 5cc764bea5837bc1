// Spectral accumulation of a block of monitor samples on the matrix cores.
//
// The running DFT of src/prismo/monitors/field.py:128-143,
//   F[f][p] += x_n[p] * exp(-i w_f t_n) * dt      for every step n,
// is, over a block of T buffered steps, one real GEMM per part:
//   Re F[f][p] += sum_n Wre[f][n] X[n][p],  Im F[f][p] += sum_n Wim[f][n] X[n][p]
// with Wre = cos(w t) dt, Wim = -sin(w t) dt (host FP64, stored in the field
// precision).  M = frequencies, N = sample columns, K = rows of the block.
//
// One wave owns a 16-column strip of X and loops over the frequencies in
// groups of 16*FT: per K-step of 4 rows each lane loads one X value
// (B operand, reused by all 2*FT products) and 2*FT phasors (A operands, a
// small table that stays in L1/L2), and issues 2*FT MFMA 16x16x4 (FP64 for
// the plain path, FP32 for the precision form).  The block's sums are added
// into the accumulators afterwards; in FP32 with Kahan-Babuska compensation
// so the error stays that of one block however many blocks a run has.
//
// Operand layout of V_MFMA_*_16X16X4 (one block): lane l holds
// A[l % 16][l / 16] and B[l / 16][l % 16]; its four D registers r = 0..3
// hold column l % 16 of rows 4 (l / 16) + r (F32) or (l / 16) + 4 r (F64),
// as measured on gfx950.
#include "pm_internal.h"

namespace pm {

template <typename T>
struct Mfma;

// row of the D element held in register r of a lane in lane group q = l / 16
template <>
struct Mfma<double> {
  using v4 = double __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int drow(int q, int r) { return q + 4 * r; }
};

template <>
struct Mfma<float> {
  using v4 = float __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int drow(int q, int r) { return 4 * q + r; }
};

template <typename T>
__device__ __forceinline__ void kb_add(T* acc, T* cmp, int64_t o, T v) {
  const T a = acc[o];
  const T t = a + v;
  const T e = (fabs(a) >= fabs(v)) ? (a - t) + v : (v - t) + a;
  acc[o] = t;
  cmp[o] += e;
}

constexpr int kDftWaves = 4;

template <typename T, int FT>
__global__ __launch_bounds__(64 * kDftWaves) void k_dft_mfma(
    const KDft d, const T* __restrict__ X, int ncols, const T* __restrict__ Wre,
    const T* __restrict__ Wim, int64_t table_len, int64_t wcol0, int rows,
    const int64_t* __restrict__ wclock) {
  using v4 = typename Mfma<T>::v4;
  // in a captured step loop the block's first step is read from the device
  // clock (wcol0 then holds the table's first step, subtracted here)
  if (wclock != nullptr) wcol0 = *wclock - wcol0;
  const int np = d.col1 - d.col0;
  const int nf = d.f1 - d.f0;
  const int lane = threadIdx.x & 63;
  const int strip = blockIdx.x * kDftWaves + (int)(threadIdx.x >> 6);
  const int p0 = strip * 16;
  if (p0 >= np) return;  // whole wave leaves together
  const int lr = lane & 15;
  const int lq = lane >> 4;
  const int p = p0 + lr;
  const bool pok = p < np;
  const T* xcol = X + d.col0 + (pok ? p : 0);
  const int64_t wbase = (int64_t)d.f0 * table_len + wcol0;

  for (int fb = 0; fb < nf; fb += 16 * FT) {
    v4 cre[FT], cim[FT];
#pragma unroll
    for (int t = 0; t < FT; ++t) {
      cre[t] = v4{0, 0, 0, 0};
      cim[t] = v4{0, 0, 0, 0};
    }
    for (int n0 = 0; n0 < rows; n0 += 4) {
      const int n = n0 + lq;
      const bool nok = n < rows;
      const T b = (pok && nok) ? xcol[(int64_t)n * ncols] : (T)0;
#pragma unroll
      for (int t = 0; t < FT; ++t) {
        const int f = fb + 16 * t + lr;
        T ar = 0, ai = 0;
        if (nok && f < nf) {
          const int64_t w = wbase + (int64_t)f * table_len + n;
          ar = Wre[w];
          ai = Wim[w];
        }
        cre[t] = Mfma<T>::mma(ar, b, cre[t]);
        cim[t] = Mfma<T>::mma(ai, b, cim[t]);
      }
    }
    if (!pok) continue;
    T* accr = (T*)d.acc_re;
    T* acci = (T*)d.acc_im;
#pragma unroll
    for (int t = 0; t < FT; ++t) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int f = fb + 16 * t + Mfma<T>::drow(lq, r);
        if (f >= nf) continue;
        const int64_t o = (int64_t)f * np + p;
        if (d.cmp_re == nullptr) {
          accr[o] += cre[t][r];
          acci[o] += cim[t][r];
        } else {
          kb_add<T>(accr, (T*)d.cmp_re, o, cre[t][r]);
          kb_add<T>(acci, (T*)d.cmp_im, o, cim[t][r]);
        }
      }
    }
  }
}

template <typename T>
static void launch_t(const KDft& d, const void* X, int ncols, const void* W_re,
                     const void* W_im, int64_t table_len, int64_t wcol0, int rows,
                     const int64_t* wclock, hipStream_t s) {
  const int np = d.col1 - d.col0, nf = d.f1 - d.f0;
  const int strips = (np + 15) / 16;
  dim3 grid((strips + kDftWaves - 1) / kDftWaves);
  dim3 block(64 * kDftWaves);
  // frequency groups of 64 (FT = 4) unless the group is small
  if (nf > 32)
    hipLaunchKernelGGL((k_dft_mfma<T, 4>), grid, block, 0, s, d, (const T*)X, ncols,
                       (const T*)W_re, (const T*)W_im, table_len, wcol0, rows, wclock);
  else if (nf > 16)
    hipLaunchKernelGGL((k_dft_mfma<T, 2>), grid, block, 0, s, d, (const T*)X, ncols,
                       (const T*)W_re, (const T*)W_im, table_len, wcol0, rows, wclock);
  else
    hipLaunchKernelGGL((k_dft_mfma<T, 1>), grid, block, 0, s, d, (const T*)X, ncols,
                       (const T*)W_re, (const T*)W_im, table_len, wcol0, rows, wclock);
}

void launch_dft(const KDft& d, const void* X, int ncols, const void* W_re, const void* W_im,
                int64_t table_len, int64_t wcol0, int rows, int precision, hipStream_t s,
                const int64_t* wclock) {
  const int np = d.col1 - d.col0, nf = d.f1 - d.f0;
  if (np <= 0 || nf <= 0 || rows <= 0) return;
  if (precision == PM_FP64)
    launch_t<double>(d, X, ncols, W_re, W_im, table_len, wcol0, rows, wclock, s);
  else
    launch_t<float>(d, X, ncols, W_re, W_im, table_len, wcol0, rows, wclock, s);
}

}  // namespace pm
