// Reference-arithmetic kernels (scheme PM_SCHEME_COMPAT), FP64.
//
// Each update is the expression of src/prismo/core/solver.py evaluated in the
// same order with the same IEEE operations: differences divided (not
// multiplied) by the spacing, then `da*H - db*(c1 - c2)` or `ca*E + cb*(...)`.
// This file is compiled with -ffp-contract=off so no multiply-add is fused,
// which keeps every field value bit-identical to the NumPy reference.  The
// averaged coefficients (solver.py:458-533) arrive precomputed per component
// point; averaging a constant array once or every step gives the same bits.
//
// Ranges are the reference's slice extents, with shapes (3-D)
//   Ex (nx, ny-1, nz-1)  Ey (nx-1, ny, nz-1)  Ez (nx-1, ny-1, nz)
//   Hx (nx-1, ny, nz)    Hy (nx, ny-1, nz)    Hz (nx, ny, nz-1):
//   solver.py:191-201   Hx[:, :ny-2, :nz-2]
//   solver.py:215-229   Hy[:nx-2, :, :nz-2]
//   solver.py:239-253   Hz[:nx-2, :ny-2, :]
//   solver.py:265-309   Ex, Ey, Ez over their full extents.
// 2-D (solver.py:311-456) with the gates of solver.py:321 and :367:
//   TM  Hx[:, :ny-2] = da Hx + db (Ez[:, j+1] - Ez[:, j]) / dy
//       Hy[:nx-2, :] = da Hy - db (Ez[i+1] - Ez[i]) / dx      if max|Ez| > 0
//   TE  Hz = da Hz - db (cx - cy), cx = 0 for i >= nx-2, cy = 0 for j >= ny-2
//                                         if max|Ex| > 0 or max|Ey| > 0
//   E   Ez = ca Ez + cb ((Hy[i+1]-Hy[i])/dx - (Hx[:, j+1]-Hx[:, j])/dy)
//       Ex = ca Ex + cb (Hz[:, j+1]-Hz[:, j])/dy ; Ey = ca Ey - cb (Hz[i+1]-Hz[i])/dx
// np.max of an array holding NaN is NaN and NaN > 0 is false, so a gate is
// "no NaN and some non-zero".  The TE E update is skipped only when Ex, Ey
// and Hz are all exactly zero, where it would write +0 back (same bits).
#include <algorithm>

#include "pm_internal.h"

namespace pm {

// flags[2a] = some element non-zero and not NaN, flags[2a+1] = some NaN,
// for a = 0: Ez, 1: Ex, 2: Ey, 3: Hz (2-D plane, padded layout), computed
// before every H update into one of two sets of eight words that steps
// alternate: the H kernel of a step clears the other set, which the next
// step's flags kernel fills (both sets start cleared at configure).
__global__ void k_flags_clear(int* f) {
  if (threadIdx.x < 16) f[threadIdx.x] = 0;
}

constexpr int kFlagRows = 4;

// One column i per thread and kFlagRows rows per workgroup (the loads of a
// thread have no loop-carried address, so they overlap); a wave ORs a flag
// only when its cells set it and it was not yet seen set (every wave
// ORing into the same eight words serialises the kernel; a stale read only
// costs an atomic that changes nothing).
__global__ __launch_bounds__(256) void k_flags2d(const KGrid g) {
  const double* A[4] = {(const double*)g.E[2], (const double*)g.E[0],
                        (const double*)g.E[1], (const double*)g.H[2]};
  const int ex[4] = {g.nx - 1, g.nx, g.nx - 1, g.nx};
  const int ey[4] = {g.ny - 1, g.ny - 1, g.ny, g.ny};
  unsigned bits = 0;  // bit 2a: non-zero and not NaN, bit 2a+1: NaN
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j0 = blockIdx.y * kFlagRows;
#pragma unroll
  for (int r = 0; r < kFlagRows; ++r) {
    const int j = j0 + r;
    const int64_t off = (int64_t)i + (int64_t)j * g.sy;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      if (A[a] == nullptr || i >= ex[a] || j >= ey[a]) continue;
      const double v = A[a][off];
      if (v != v) bits |= 2u << (2 * a);
      else if (v != 0.0) bits |= 1u << (2 * a);
    }
  }
  for (int m = 32; m >= 1; m >>= 1) bits |= (unsigned)__shfl_xor((int)bits, m);
  if ((threadIdx.x % warpSize) == 0 && bits) {
    volatile const int* fv = g.flags;
#pragma unroll
    for (int b = 0; b < 8; ++b)
      if (((bits >> b) & 1u) && !fv[b]) atomicOr(&g.flags[b], 1);
  }
}

template <int CK>
__device__ __forceinline__ double hco_a(const KGrid& g, int c, int64_t off) {
  if constexpr (CK == PM_COEFF_FULL) return ((const double*)g.da[c])[off];
  return 1.0;
}
template <int CK>
__device__ __forceinline__ double hco_b(const KGrid& g, int c, int64_t off) {
  if constexpr (CK == PM_COEFF_FULL) return ((const double*)g.db[c])[off];
  return g.db_const;
}
template <int CK>
__device__ __forceinline__ double eco_a(const KGrid& g, int c, int64_t off) {
  if constexpr (CK == PM_COEFF_FULL) return ((const double*)g.ca[c])[off];
  return 1.0;
}
__device__ __forceinline__ double eco_b(const KGrid& g, int c, int64_t off) {
  return ((const double*)g.cb[c])[off];
}

// ------------------------------------------------------------------ 2-D
template <int CK>
__global__ __launch_bounds__(kBX* kBY) void k_compat_h2d(const KGrid g) {
  const int i = blockIdx.x * kBX + threadIdx.x;
  const int j = blockIdx.y * kBY + threadIdx.y;
  // the next step's flag set (read by nothing in this step)
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.y == 0 && threadIdx.x < 8)
    g.flags_next[threadIdx.x] = 0;
  if (i >= g.nx || j >= g.ny) return;
  const int* f = g.flags;
  const bool g_tm = f[0] && !f[1];
  const bool g_te = (f[2] && !f[3]) || (f[4] && !f[5]);
  const int64_t off = (int64_t)i + (int64_t)j * g.sy;
  const double* Ez = (const double*)g.E[2];
  if (g_tm) {
    if (i < g.nx - 1 && j < g.ny - 2) {
      double* Hx = (double*)g.H[0];
      const double c = (Ez[off + g.sy] - Ez[off]) / g.dy;
      Hx[off] = hco_a<CK>(g, 0, off) * Hx[off] + hco_b<CK>(g, 0, off) * c;
    }
    if (i < g.nx - 2 && j < g.ny - 1) {
      double* Hy = (double*)g.H[1];
      const double c = (Ez[off + 1] - Ez[off]) / g.dx;
      Hy[off] = hco_a<CK>(g, 1, off) * Hy[off] - hco_b<CK>(g, 1, off) * c;
    }
  }
  if (g_te) {
    const double* Ex = (const double*)g.E[0];
    const double* Ey = (const double*)g.E[1];
    double* Hz = (double*)g.H[2];
    const double cx = (i < g.nx - 2) ? (Ey[off + 1] - Ey[off]) / g.dx : 0.0;
    const double cy = (j < g.ny - 2) ? (Ex[off + g.sy] - Ex[off]) / g.dy : 0.0;
    Hz[off] = hco_a<CK>(g, 2, off) * Hz[off] - hco_b<CK>(g, 2, off) * (cx - cy);
  }
}

template <int CK>
__global__ __launch_bounds__(kBX* kBY) void k_compat_e2d(const KGrid g) {
  const int i = blockIdx.x * kBX + threadIdx.x;
  const int j = blockIdx.y * kBY + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const int* f = g.flags;
  const bool te_any = f[2] | f[3] | f[4] | f[5] | f[6] | f[7];
  const int64_t off = (int64_t)i + (int64_t)j * g.sy;
  const double* Hx = (const double*)g.H[0];
  const double* Hy = (const double*)g.H[1];
  if (i < g.nx - 1 && j < g.ny - 1) {
    double* Ez = (double*)g.E[2];
    const double cx = (Hy[off + 1] - Hy[off]) / g.dx;
    const double cy = (Hx[off + g.sy] - Hx[off]) / g.dy;
    Ez[off] = eco_a<CK>(g, 2, off) * Ez[off] + eco_b(g, 2, off) * (cx - cy);
  }
  if (te_any) {
    const double* Hz = (const double*)g.H[2];
    if (j < g.ny - 1) {
      double* Ex = (double*)g.E[0];
      const double c = (Hz[off + g.sy] - Hz[off]) / g.dy;
      Ex[off] = eco_a<CK>(g, 0, off) * Ex[off] + eco_b(g, 0, off) * c;
    }
    if (i < g.nx - 1) {
      double* Ey = (double*)g.E[1];
      const double c = (Hz[off + 1] - Hz[off]) / g.dx;
      Ey[off] = eco_a<CK>(g, 1, off) * Ey[off] - eco_b(g, 1, off) * c;
    }
  }
}

// ------------------------------------------------------------------ 3-D
template <int CK>
__global__ __launch_bounds__(kBX* kBY) void k_compat_h3d(const KGrid g) {
  const int i = blockIdx.x * kBX + threadIdx.x;
  const int j = blockIdx.y * kBY + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  const int64_t off = (int64_t)i + (int64_t)j * g.sy + (int64_t)k * g.sz;
  const double* Ex = (const double*)g.E[0];
  const double* Ey = (const double*)g.E[1];
  const double* Ez = (const double*)g.E[2];
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (i < nx - 1 && j < ny - 2 && k < nz - 2) {
    double* H = (double*)g.H[0];
    const double c1 = (Ez[off + g.sy] - Ez[off]) / g.dy;
    const double c2 = (Ey[off + g.sz] - Ey[off]) / g.dz;
    H[off] = hco_a<CK>(g, 0, off) * H[off] - hco_b<CK>(g, 0, off) * (c1 - c2);
  }
  if (i < nx - 2 && j < ny - 1 && k < nz - 2) {
    double* H = (double*)g.H[1];
    const double c1 = (Ex[off + g.sz] - Ex[off]) / g.dz;
    const double c2 = (Ez[off + 1] - Ez[off]) / g.dx;
    H[off] = hco_a<CK>(g, 1, off) * H[off] - hco_b<CK>(g, 1, off) * (c1 - c2);
  }
  if (i < nx - 2 && j < ny - 2 && k < nz - 1) {
    double* H = (double*)g.H[2];
    const double c1 = (Ey[off + 1] - Ey[off]) / g.dx;
    const double c2 = (Ex[off + g.sy] - Ex[off]) / g.dy;
    H[off] = hco_a<CK>(g, 2, off) * H[off] - hco_b<CK>(g, 2, off) * (c1 - c2);
  }
}

template <int CK>
__global__ __launch_bounds__(kBX* kBY) void k_compat_e3d(const KGrid g) {
  const int i = blockIdx.x * kBX + threadIdx.x;
  const int j = blockIdx.y * kBY + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  const int64_t off = (int64_t)i + (int64_t)j * g.sy + (int64_t)k * g.sz;
  const double* Hx = (const double*)g.H[0];
  const double* Hy = (const double*)g.H[1];
  const double* Hz = (const double*)g.H[2];
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (j < ny - 1 && k < nz - 1) {
    double* E = (double*)g.E[0];
    const double c1 = (Hz[off + g.sy] - Hz[off]) / g.dy;
    const double c2 = (Hy[off + g.sz] - Hy[off]) / g.dz;
    E[off] = eco_a<CK>(g, 0, off) * E[off] + eco_b(g, 0, off) * (c1 - c2);
  }
  if (i < nx - 1 && k < nz - 1) {
    double* E = (double*)g.E[1];
    const double c1 = (Hx[off + g.sz] - Hx[off]) / g.dz;
    const double c2 = (Hz[off + 1] - Hz[off]) / g.dx;
    E[off] = eco_a<CK>(g, 1, off) * E[off] + eco_b(g, 1, off) * (c1 - c2);
  }
  if (i < nx - 1 && j < ny - 1) {
    double* E = (double*)g.E[2];
    const double c1 = (Hy[off + 1] - Hy[off]) / g.dx;
    const double c2 = (Hx[off + g.sy] - Hx[off]) / g.dy;
    E[off] = eco_a<CK>(g, 2, off) * E[off] + eco_b(g, 2, off) * (c1 - c2);
  }
}

void launch_compat_flags_clear(int* flags, hipStream_t s) {
  hipLaunchKernelGGL(k_flags_clear, dim3(1), dim3(64), 0, s, flags);
}

void launch_compat_flags2d(const KGrid& g, hipStream_t s) {
  const dim3 grid((g.nx + 255) / 256, (g.ny + kFlagRows - 1) / kFlagRows);
  hipLaunchKernelGGL(k_flags2d, grid, dim3(256), 0, s, g);
}

void launch_compat_h(const KGrid& g, int ck, hipStream_t s) {
  dim3 block(kBX, kBY);
  dim3 grid((g.nx + kBX - 1) / kBX, (g.ny + kBY - 1) / kBY, g.ndim == 3 ? g.nz : 1);
  if (g.ndim == 2) {
    if (ck == PM_COEFF_FULL) hipLaunchKernelGGL(k_compat_h2d<PM_COEFF_FULL>, grid, block, 0, s, g);
    else hipLaunchKernelGGL(k_compat_h2d<PM_COEFF_CB>, grid, block, 0, s, g);
  } else {
    if (ck == PM_COEFF_FULL) hipLaunchKernelGGL(k_compat_h3d<PM_COEFF_FULL>, grid, block, 0, s, g);
    else hipLaunchKernelGGL(k_compat_h3d<PM_COEFF_CB>, grid, block, 0, s, g);
  }
}

void launch_compat_e(const KGrid& g, int ck, hipStream_t s) {
  dim3 block(kBX, kBY);
  dim3 grid((g.nx + kBX - 1) / kBX, (g.ny + kBY - 1) / kBY, g.ndim == 3 ? g.nz : 1);
  if (g.ndim == 2) {
    if (ck == PM_COEFF_FULL) hipLaunchKernelGGL(k_compat_e2d<PM_COEFF_FULL>, grid, block, 0, s, g);
    else hipLaunchKernelGGL(k_compat_e2d<PM_COEFF_CB>, grid, block, 0, s, g);
  } else {
    if (ck == PM_COEFF_FULL) hipLaunchKernelGGL(k_compat_e3d<PM_COEFF_FULL>, grid, block, 0, s, g);
    else hipLaunchKernelGGL(k_compat_e3d<PM_COEFF_CB>, grid, block, 0, s, g);
  }
}

}  // namespace pm
