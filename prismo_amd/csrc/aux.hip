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
//
// A step takes one launch for all sample boxes and one for all source
// boxes (blockIdx.y = box, box table in device memory): small grids are
// launch-bound, and a plane-wave source with three port monitors would
// otherwise take eight launches per step.  Sources whose boxes of one
// component overlap keep one launch per box, so their adds happen in the
// reference's order.
#include <algorithm>

#include "pm_internal.h"

namespace pm {

// Table row of a source: `row`, plus the step clock when one is given (a
// captured step replays with the clock holding its first step).
__device__ __forceinline__ int64_t source_row(int64_t row, const int64_t* clock) {
  return clock ? row + *clock : row;
}

template <typename T>
__global__ __launch_bounds__(256) void k_source(const KSrc s, int64_t row, int64_t sy,
                                                int64_t sz, T* F, const int64_t* clock) {
  const int ni = s.i1 - s.i0, nj = s.j1 - s.j0, nk = s.k1 - s.k0;
  const int64_t n = (int64_t)ni * nj * nk;
  const double a = s.table[source_row(row, clock)];
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

// Field pointer of component c (uniform) from six scalar arguments: an
// indexed argument array would be copied to private memory.
template <typename T>
__device__ __forceinline__ T* field_of(int c, T* f0, T* f1, T* f2, T* f3, T* f4, T* f5) {
  return c == 0 ? f0 : c == 1 ? f1 : c == 2 ? f2 : c == 3 ? f3 : c == 4 ? f4 : f5;
}

template <typename T>
__global__ __launch_bounds__(256) void k_source_all(const KSrc* tab, int64_t row, int64_t sy,
                                                    int64_t sz, T* f0, T* f1, T* f2, T* f3,
                                                    T* f4, T* f5, const int64_t* clock) {
  const KSrc s = tab[blockIdx.y];  // a kernel argument: global memory
  const int ni = s.i1 - s.i0, nj = s.j1 - s.j0, nk = s.k1 - s.k0;
  const int64_t n = (int64_t)ni * nj * nk;
  T* F = field_of<T>(s.comp, f0, f1, f2, f3, f4, f5);
  const double a = ((gptr<const double>)s.table)[source_row(row, clock)];
  const gptr<const double> prof = (gptr<const double>)s.profile;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(c % ni);
    const int64_t r = c / ni;
    const int j = (int)(r % nj), k = (int)(r / nj);
    double v = a;
    if (s.kind == 1) {
      v = a * prof[c];
      if (s.divisor != 0.0) v = v / s.divisor;
    }
    const int64_t off = (int64_t)(s.i0 + i) + (int64_t)(s.j0 + j) * sy + (int64_t)(s.k0 + k) * sz;
    F[off] = F[off] + (T)v;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_sample_all(const KSbox* tab, int64_t sy, int64_t sz,
                                                    const T* f0, const T* f1, const T* f2,
                                                    const T* f3, const T* f4, const T* f5,
                                                    T* Xrow) {
  const KSbox b = tab[blockIdx.y];
  const int ni = b.i1 - b.i0, nj = b.j1 - b.j0, nk = b.k1 - b.k0;
  const int64_t n = (int64_t)ni * nj * nk;
  const T* F = field_of<const T>(b.comp, f0, f1, f2, f3, f4, f5);
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(c % ni);
    const int64_t r = c / ni;
    const int j = (int)(r % nj), k = (int)(r / nj);
    const int64_t off = (int64_t)(b.i0 + i) + (int64_t)(b.j0 + j) * sy + (int64_t)(b.k0 + k) * sz;
    Xrow[b.col0 + c] = F[off];
  }
}

// Sources (blockIdx.y < n_src, as k_source_all) and samples (the rest, as
// k_sample_all) of one step in one launch; only for tables where no sample
// box meets a source box of its component.
template <typename T>
__global__ __launch_bounds__(256) void k_source_sample(const KSrc* stab, int n_src,
                                                       const KSbox* btab, int64_t row,
                                                       int64_t sy, int64_t sz, T* f0, T* f1,
                                                       T* f2, T* f3, T* f4, T* f5, T* Xrow,
                                                       const int64_t* clock) {
  const int y = blockIdx.y;
  if (y < n_src) {
    const KSrc s = stab[y];
    const int ni = s.i1 - s.i0, nj = s.j1 - s.j0, nk = s.k1 - s.k0;
    const int64_t n = (int64_t)ni * nj * nk;
    T* F = field_of<T>(s.comp, f0, f1, f2, f3, f4, f5);
    const double a = ((gptr<const double>)s.table)[source_row(row, clock)];
    const gptr<const double> prof = (gptr<const double>)s.profile;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
      const int i = (int)(c % ni);
      const int64_t r = c / ni;
      const int j = (int)(r % nj), k = (int)(r / nj);
      double v = a;
      if (s.kind == 1) {
        v = a * prof[c];
        if (s.divisor != 0.0) v = v / s.divisor;
      }
      const int64_t off =
          (int64_t)(s.i0 + i) + (int64_t)(s.j0 + j) * sy + (int64_t)(s.k0 + k) * sz;
      F[off] = F[off] + (T)v;
    }
    return;
  }
  const KSbox b = btab[y - n_src];
  const int ni = b.i1 - b.i0, nj = b.j1 - b.j0, nk = b.k1 - b.k0;
  const int64_t n = (int64_t)ni * nj * nk;
  const T* F = field_of<T>(b.comp, f0, f1, f2, f3, f4, f5);
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

template <typename T>
static void launch_source_all(const KAux& a, int64_t row, hipStream_t s, const int64_t* clock) {
  T* const* F = (T* const*)a.F;
  hipLaunchKernelGGL(k_source_all<T>, dim3(blocks_for(a.src_max), a.n_src), dim3(256), 0, s,
                     a.src_dev, row, a.sy, a.sz, F[0], F[1], F[2], F[3], F[4], F[5], clock);
}

template <typename T>
static void launch_sample_all(const KAux& a, int64_t row, hipStream_t s) {
  T* const* F = (T* const*)a.F;
  hipLaunchKernelGGL(k_sample_all<T>, dim3(blocks_for(a.sbox_max), a.n_sbox), dim3(256), 0, s,
                     a.sbox_dev, a.sy, a.sz, (const T*)F[0], (const T*)F[1], (const T*)F[2],
                     (const T*)F[3], (const T*)F[4], (const T*)F[5],
                     (T*)a.X + row * (int64_t)a.ncols);
}

void launch_sources(const KAux& a, int64_t step, hipStream_t s, const int64_t* clock) {
  const int64_t row = step - a.table_step0;
  if (a.n_src > 1 && a.src_merge && a.src_max > 0) {
    if (a.precision == PM_FP64)
      launch_source_all<double>(a, row, s, clock);
    else
      launch_source_all<float>(a, row, s, clock);
    return;
  }
  for (int t = 0; t < a.n_src; ++t) {
    const KSrc& src = a.src[t];
    const int64_t n = (int64_t)(src.i1 - src.i0) * (src.j1 - src.j0) * (src.k1 - src.k0);
    if (n <= 0) continue;
    if (a.precision == PM_FP64)
      hipLaunchKernelGGL(k_source<double>, dim3(blocks_for(n)), dim3(256), 0, s, src, row,
                         a.sy, a.sz, (double*)a.F[src.comp], clock);
    else
      hipLaunchKernelGGL(k_source<float>, dim3(blocks_for(n)), dim3(256), 0, s, src, row,
                         a.sy, a.sz, (float*)a.F[src.comp], clock);
  }
}

void launch_sources_sample(const KAux& a, int64_t step, int64_t row, hipStream_t s,
                           const int64_t* clock) {
  const int64_t srow = step - a.table_step0;
  const int64_t xrow = row % a.sample_T;
  const dim3 grid(blocks_for(std::max(a.src_max, a.sbox_max)), a.n_src + a.n_sbox);
  if (a.precision == PM_FP64) {
    double* const* F = (double* const*)a.F;
    hipLaunchKernelGGL(k_source_sample<double>, grid, dim3(256), 0, s, a.src_dev, a.n_src,
                       a.sbox_dev, srow, a.sy, a.sz, F[0], F[1], F[2], F[3], F[4], F[5],
                       (double*)a.X + xrow * (int64_t)a.ncols, clock);
  } else {
    float* const* F = (float* const*)a.F;
    hipLaunchKernelGGL(k_source_sample<float>, grid, dim3(256), 0, s, a.src_dev, a.n_src,
                       a.sbox_dev, srow, a.sy, a.sz, F[0], F[1], F[2], F[3], F[4], F[5],
                       (float*)a.X + xrow * (int64_t)a.ncols, clock);
  }
}

void launch_sample(const KAux& a, int64_t step, hipStream_t s) {
  const int64_t row = step % a.sample_T;
  if (a.n_sbox > 1 && a.sbox_max > 0) {
    if (a.precision == PM_FP64)
      launch_sample_all<double>(a, row, s);
    else
      launch_sample_all<float>(a, row, s);
    return;
  }
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

__global__ void k_step_clock(int64_t* clock, int64_t v, int add) {
  if (threadIdx.x == 0) *clock = add ? *clock + v : v;
}

void launch_clock(int64_t* clock, int64_t v, bool add, hipStream_t s) {
  hipLaunchKernelGGL(k_step_clock, dim3(1), dim3(64), 0, s, clock, v, add ? 1 : 0);
}

}  // namespace pm
