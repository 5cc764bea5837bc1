// Source injection and monitor sampling (the spectral accumulation of the
// sampled blocks is in dft_mfma.hip).
//
// Sources: field[c] += value on a box (src/prismo/sources/point.py:66-73,
// plane_wave.py:206-221, gaussian.py:270-298).  For FP32 fields the value is
// rounded to FP32 first and added in FP32, which is what NumPy does for a
// float32 array plus a Python float.
//
// Sampling copies each monitored box (the region of
// src/prismo/monitors/base.py:58-134) into one row of the sample matrix X;
// the spectra are then one dense product per flushed block
// (src/prismo/monitors/field.py:138-143 summed over a block of steps).
#include "pm_internal.h"

namespace pm {

template <typename T>
__global__ __launch_bounds__(256) void k_source(const KSrc s, int64_t row, int64_t sy,
                                                int64_t sz, T* F) {
  const int ni = s.i1 - s.i0, nj = s.j1 - s.j0, nk = s.k1 - s.k0;
  const int64_t n = (int64_t)ni * nj * nk;
  const double a = s.table[row];
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(c % ni);
    const int64_t r = c / ni;
    const int j = (int)(r % nj), k = (int)(r / nj);
    double v = a;
    if (s.kind == 1) {
      v = a * s.profile[c];
      if (s.divisor != 0.0) v = v / s.divisor;
    }
    const int64_t off = (int64_t)(s.i0 + i) + (int64_t)(s.j0 + j) * sy + (int64_t)(s.k0 + k) * sz;
    F[off] = F[off] + (T)v;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_sample(const KSbox b, int64_t sy, int64_t sz,
                                                const T* F, T* Xrow) {
  const int ni = b.i1 - b.i0, nj = b.j1 - b.j0, nk = b.k1 - b.k0;
  const int64_t n = (int64_t)ni * nj * nk;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(c % ni);
    const int64_t r = c / ni;
    const int j = (int)(r % nj), k = (int)(r / nj);
    const int64_t off = (int64_t)(b.i0 + i) + (int64_t)(b.j0 + j) * sy + (int64_t)(b.k0 + k) * sz;
    Xrow[b.col0 + c] = F[off];
  }
}

static int blocks_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > 2048) b = 2048;
  return (int)b;
}

void launch_sources(const KAux& a, int64_t step, hipStream_t s) {
  const int64_t row = step - a.table_step0;
  for (int t = 0; t < a.n_src; ++t) {
    const KSrc& src = a.src[t];
    const int64_t n = (int64_t)(src.i1 - src.i0) * (src.j1 - src.j0) * (src.k1 - src.k0);
    if (n <= 0) continue;
    if (a.precision == PM_FP64)
      hipLaunchKernelGGL(k_source<double>, dim3(blocks_for(n)), dim3(256), 0, s, src, row,
                         a.sy, a.sz, (double*)a.F[src.comp]);
    else
      hipLaunchKernelGGL(k_source<float>, dim3(blocks_for(n)), dim3(256), 0, s, src, row,
                         a.sy, a.sz, (float*)a.F[src.comp]);
  }
}

void launch_sample(const KAux& a, int64_t step, hipStream_t s) {
  const int64_t row = step % a.sample_T;
  for (int t = 0; t < a.n_sbox; ++t) {
    const KSbox& b = a.sbox[t];
    const int64_t n = (int64_t)(b.i1 - b.i0) * (b.j1 - b.j0) * (b.k1 - b.k0);
    if (n <= 0) continue;
    if (a.precision == PM_FP64)
      hipLaunchKernelGGL(k_sample<double>, dim3(blocks_for(n)), dim3(256), 0, s, b, a.sy,
                         a.sz, (const double*)a.F[b.comp],
                         (double*)a.X + row * (int64_t)a.ncols);
    else
      hipLaunchKernelGGL(k_sample<float>, dim3(blocks_for(n)), dim3(256), 0, s, b, a.sy,
                         a.sz, (const float*)a.F[b.comp],
                         (float*)a.X + row * (int64_t)a.ncols);
  }
}

}  // namespace pm
