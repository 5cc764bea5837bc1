// Textbook Yee leapfrog kernels (scheme PM_SCHEME_YEE) for gfx950.
//
// Staggering is the reference's (src/prismo/core/solver.py:173-175,261-263):
// E on cell faces, H on cell edges.  The E update is the reference's
// (solver.py:255-309, forward differences of H); the H update takes backward
// differences of E, which is the consistent Yee pairing (the reference's
// forward H differences at solver.py:178-253 are one cell off and unstable).
// E samples that fall outside the grid read as zero: a PEC wall half a cell
// beyond the outermost H nodes.
//
// CPML (Roden & Gedney 2000): for every derivative term d along axis a inside
// an a-slab,  psi <- b psi + c d,  term = d / kappa + psi.  psi lives only in
// the slabs (pm_plan.psi_*).  ADE regions replace the E update of their cells
// by the coupled polarisation update (see ade_update below).
//
// Layout and traffic: one thread per (i, j) column, x-fastest rows of 64 lanes
// (one wave64), marching `kchunk` planes in z and carrying the k-1 (H kernel)
// or k+1 (E kernel) plane in registers, so each field element is read from
// memory once per kernel apart from the in-plane +-1 neighbours, which hit
// the L1/L2 lines the neighbouring lanes and rows just fetched.
#include "pm_internal.h"

namespace pm {

__device__ __forceinline__ int slab_of(int p, int n, int tlo, int thi) {
  if (p < tlo) return p;
  if (p >= n - thi) return tlo + (p - (n - thi));
  return -1;
}

// Offset of (i, j, k) with slab index s along axis a in a psi slab array.
__device__ __forceinline__ int64_t slab_off(const KGrid& g, int a, int s, int i,
                                            int j, int k) {
  if (a == 0) {
    const int S = g.pml[0] + g.pml[1];
    return (int64_t)s + (int64_t)S * ((int64_t)j + (int64_t)g.ny * k);
  } else if (a == 1) {
    const int S = g.pml[2] + g.pml[3];
    return (int64_t)i + (int64_t)g.nx * ((int64_t)s + (int64_t)S * k);
  }
  return (int64_t)i + (int64_t)g.nx * ((int64_t)j + (int64_t)g.ny * s);
}

// CPML-stretched derivative term.  E == true: E-update term (profiles at half
// nodes); false: H-update term (profiles at nodes).
template <typename T, bool E>
__device__ __forceinline__ T cpml_term(const KGrid& g, T d, int a, int q, int p,
                                       int s, int i, int j, int k) {
  const T* b = (const T*)(E ? g.pml_be[a] : g.pml_bh[a]);
  const T* c = (const T*)(E ? g.pml_ce[a] : g.pml_ch[a]);
  const T* ik = (const T*)(E ? g.pml_ike[a] : g.pml_ikh[a]);
  T* psi = (T*)(E ? g.psi_e[a][q] : g.psi_h[a][q]) + slab_off(g, a, s, i, j, k);
  const T ps = b[p] * (*psi) + c[p] * d;
  *psi = ps;
  return d * ik[p] + ps;
}

template <typename T, int CK>
__device__ __forceinline__ T load_cb(const KGrid& g, int c, int64_t off) {
  if constexpr (CK == PM_COEFF_CB_F16) {
    return (T)g.cb_scale * (T)__half2float(((const __half*)g.cb[c])[off]);
  } else {
    return ((const T*)g.cb[c])[off];
  }
}

// Dispersive E update: the first box of the table holding (comp, i, j, k)
// replaces the plain update (ade_cell in pm_internal.h).  Returns true if the
// cell is dispersive.
template <typename T>
__device__ __forceinline__ bool ade_update(const KGrid& g, int comp, int i, int j,
                                           int k, T curl, T& e) {
  // the box table lives in device memory, so a rolled loop indexes it
  // without copying kernel arguments to private memory
#pragma unroll 1
  for (int b = 0; b < g.n_ade; ++b) {
    const KAde& a = g.ade[b];
    if (a.comp != comp || i < a.i0 || i >= a.i1 || j < a.j0 || j >= a.j1 ||
        k < a.k0 || k >= a.k1)
      continue;
    const int64_t l = (int64_t)(i - a.i0) +
                      (int64_t)(a.i1 - a.i0) *
                          ((int64_t)(j - a.j0) + (int64_t)(a.j1 - a.j0) * (k - a.k0));
    e = ade_cell<T>(a, l, curl, e);
    return true;
  }
  return false;
}

template <typename T, int CK>
__device__ __forceinline__ void e_apply(const KGrid& g, int c, int64_t off, int i,
                                        int j, int k, T curl) {
  T* F = (T*)g.E[c];
  T e = F[off];
  if (g.n_ade && ade_update<T>(g, c, i, j, k, curl, e)) {
    F[off] = e;
    return;
  }
  const T cb = load_cb<T, CK>(g, c, off);
  if constexpr (CK == PM_COEFF_FULL) {
    F[off] = ((const T*)g.ca[c])[off] * e + cb * curl;
  } else {
    F[off] = e + cb * curl;
  }
}

template <typename T, int CK>
__device__ __forceinline__ void h_apply(const KGrid& g, int c, int64_t off, T curl) {
  T* F = (T*)g.H[c];
  const T h = F[off];
  if constexpr (CK == PM_COEFF_FULL) {
    F[off] = ((const T*)g.da[c])[off] * h - ((const T*)g.db[c])[off] * curl;
  } else {
    F[off] = h - (T)g.db_const * curl;
  }
}

// ---------------------------------------------------------------- H update
template <typename T, int CK, int NDIM>
__global__ __launch_bounds__(kBX* kBY) void k_yee_h(const KGrid g, int kchunk) {
  const int i = blockIdx.x * kBX + threadIdx.x;
  const int j = blockIdx.y * kBY + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const int kb = blockIdx.z * kchunk;
  const int ke = min(kb + kchunk, g.nz);
  const T* Ex = (const T*)g.E[0];
  const T* Ey = (const T*)g.E[1];
  const T* Ez = (const T*)g.E[2];
  const T idx = (T)g.idx, idy = (T)g.idy, idz = (T)g.idz;
  const bool hx_ok = i < g.nx - 1;  // Hx (nx-1, ny, nz)
  const bool hy_ok = j < g.ny - 1;  // Hy (nx, ny-1, nz)
  const int sxs = g.has_pml ? slab_of(i, g.nx, g.pml[0], g.pml[1]) : -1;
  const int sys = g.has_pml ? slab_of(j, g.ny, g.pml[2], g.pml[3]) : -1;
  int64_t off = (int64_t)i + (int64_t)j * g.sy + (int64_t)kb * g.sz;
  T ex_km1 = 0, ey_km1 = 0;
  if (NDIM == 3 && kb > 0) {
    ex_km1 = Ex[off - g.sz];
    ey_km1 = Ey[off - g.sz];
  }
  for (int k = kb; k < ke; ++k, off += g.sz) {
    const T ez = Ez[off];
    const T ez_jm1 = j > 0 ? Ez[off - g.sy] : (T)0;
    const T ez_im1 = i > 0 ? Ez[off - 1] : (T)0;
    T ex = 0, ey = 0;
    if (NDIM == 3) {
      ex = Ex[off];
      ey = Ey[off];
    }
    const int szs = (NDIM == 3 && g.has_pml) ? slab_of(k, g.nz, g.pml[4], g.pml[5]) : -1;
    if (hx_ok) {  // curl_x E = dEz/dy - dEy/dz
      T d1 = (ez - ez_jm1) * idy;
      T d2 = NDIM == 3 ? (ey - ey_km1) * idz : (T)0;
      if (sys >= 0) d1 = cpml_term<T, false>(g, d1, 1, 1, j, sys, i, j, k);
      if (NDIM == 3 && szs >= 0) d2 = cpml_term<T, false>(g, d2, 2, 0, k, szs, i, j, k);
      h_apply<T, CK>(g, 0, off, d1 - d2);
    }
    if (hy_ok) {  // curl_y E = dEx/dz - dEz/dx
      T d1 = NDIM == 3 ? (ex - ex_km1) * idz : (T)0;
      T d2 = (ez - ez_im1) * idx;
      if (NDIM == 3 && szs >= 0) d1 = cpml_term<T, false>(g, d1, 2, 1, k, szs, i, j, k);
      if (sxs >= 0) d2 = cpml_term<T, false>(g, d2, 0, 0, i, sxs, i, j, k);
      h_apply<T, CK>(g, 1, off, d1 - d2);
    }
    if (NDIM == 3) {
      if (k < g.nz - 1) {  // Hz (nx, ny, nz-1): curl_z E = dEy/dx - dEx/dy
        const T ey_im1 = i > 0 ? Ey[off - 1] : (T)0;
        const T ex_jm1 = j > 0 ? Ex[off - g.sy] : (T)0;
        T d1 = (ey - ey_im1) * idx;
        T d2 = (ex - ex_jm1) * idy;
        if (sxs >= 0) d1 = cpml_term<T, false>(g, d1, 0, 1, i, sxs, i, j, k);
        if (sys >= 0) d2 = cpml_term<T, false>(g, d2, 1, 0, j, sys, i, j, k);
        h_apply<T, CK>(g, 2, off, d1 - d2);
      }
      ex_km1 = ex;
      ey_km1 = ey;
    }
  }
}

// ---------------------------------------------------------------- E update
template <typename T, int CK, int NDIM>
__global__ __launch_bounds__(kBX* kBY) void k_yee_e(const KGrid g, int kchunk) {
  const int i = blockIdx.x * kBX + threadIdx.x;
  const int j = blockIdx.y * kBY + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const int kb = blockIdx.z * kchunk;
  const int ke = min(kb + kchunk, g.nz);
  const T* Hx = (const T*)g.H[0];
  const T* Hy = (const T*)g.H[1];
  const T* Hz = (const T*)g.H[2];
  const T idx = (T)g.idx, idy = (T)g.idy, idz = (T)g.idz;
  const bool ex_ok = j < g.ny - 1;                    // Ex (nx, ny-1, nz-1)
  const bool ey_ok = i < g.nx - 1;                    // Ey (nx-1, ny, nz-1)
  const bool ez_ok = (i < g.nx - 1) && (j < g.ny - 1);  // Ez (nx-1, ny-1, nz)
  const int sxs = g.has_pml ? slab_of(i, g.nx, g.pml[0], g.pml[1]) : -1;
  const int sys = g.has_pml ? slab_of(j, g.ny, g.pml[2], g.pml[3]) : -1;
  int64_t off = (int64_t)i + (int64_t)j * g.sy + (int64_t)kb * g.sz;
  T hx = Hx[off], hy = Hy[off];
  for (int k = kb; k < ke; ++k, off += g.sz) {
    const int szs = (NDIM == 3 && g.has_pml) ? slab_of(k, g.nz, g.pml[4], g.pml[5]) : -1;
    T hx_n = 0, hy_n = 0, hz = 0;
    if (NDIM == 3) {
      hz = Hz[off];
      if (k + 1 < g.nz) {
        hx_n = Hx[off + g.sz];
        hy_n = Hy[off + g.sz];
      }
      if (k < g.nz - 1) {
        if (ex_ok) {  // curl_x H = dHz/dy - dHy/dz
          T d1 = (Hz[off + g.sy] - hz) * idy;
          T d2 = (hy_n - hy) * idz;
          if (sys >= 0) d1 = cpml_term<T, true>(g, d1, 1, 1, j, sys, i, j, k);
          if (szs >= 0) d2 = cpml_term<T, true>(g, d2, 2, 0, k, szs, i, j, k);
          e_apply<T, CK>(g, 0, off, i, j, k, d1 - d2);
        }
        if (ey_ok) {  // curl_y H = dHx/dz - dHz/dx
          T d1 = (hx_n - hx) * idz;
          T d2 = (Hz[off + 1] - hz) * idx;
          if (szs >= 0) d1 = cpml_term<T, true>(g, d1, 2, 1, k, szs, i, j, k);
          if (sxs >= 0) d2 = cpml_term<T, true>(g, d2, 0, 0, i, sxs, i, j, k);
          e_apply<T, CK>(g, 1, off, i, j, k, d1 - d2);
        }
      }
    }
    if (ez_ok) {  // curl_z H = dHy/dx - dHx/dy
      T d1 = (Hy[off + 1] - hy) * idx;
      T d2 = (Hx[off + g.sy] - hx) * idy;
      if (sxs >= 0) d1 = cpml_term<T, true>(g, d1, 0, 1, i, sxs, i, j, k);
      if (sys >= 0) d2 = cpml_term<T, true>(g, d2, 1, 0, j, sys, i, j, k);
      e_apply<T, CK>(g, 2, off, i, j, k, d1 - d2);
    }
    if (NDIM == 3) {
      hx = hx_n;
      hy = hy_n;
    }
  }
}

static int kchunk_for(int nz) { return nz >= 64 ? 16 : (nz >= 8 ? 4 : 1); }

template <typename T, int CK>
static void launch_h_t(const KGrid& g, hipStream_t s) {
  const int kc = g.ndim == 3 ? kchunk_for(g.nz) : 1;
  dim3 grid((g.nx + kBX - 1) / kBX, (g.ny + kBY - 1) / kBY, (g.nz + kc - 1) / kc);
  dim3 block(kBX, kBY);
  if (g.ndim == 3)
    hipLaunchKernelGGL((k_yee_h<T, CK, 3>), grid, block, 0, s, g, kc);
  else
    hipLaunchKernelGGL((k_yee_h<T, CK, 2>), grid, block, 0, s, g, kc);
}

template <typename T, int CK>
static void launch_e_t(const KGrid& g, hipStream_t s) {
  const int kc = g.ndim == 3 ? kchunk_for(g.nz) : 1;
  dim3 grid((g.nx + kBX - 1) / kBX, (g.ny + kBY - 1) / kBY, (g.nz + kc - 1) / kc);
  dim3 block(kBX, kBY);
  if (g.ndim == 3)
    hipLaunchKernelGGL((k_yee_e<T, CK, 3>), grid, block, 0, s, g, kc);
  else
    hipLaunchKernelGGL((k_yee_e<T, CK, 2>), grid, block, 0, s, g, kc);
}

void launch_yee_h(const KGrid& g, int precision, int ck, hipStream_t s) {
  if (precision == PM_FP64) {
    if (ck == PM_COEFF_FULL) launch_h_t<double, PM_COEFF_FULL>(g, s);
    else launch_h_t<double, PM_COEFF_CB>(g, s);
  } else {
    if (ck == PM_COEFF_FULL) launch_h_t<float, PM_COEFF_FULL>(g, s);
    else if (ck == PM_COEFF_CB) launch_h_t<float, PM_COEFF_CB>(g, s);
    else launch_h_t<float, PM_COEFF_CB_F16>(g, s);
  }
}

void launch_yee_e(const KGrid& g, int precision, int ck, hipStream_t s) {
  if (precision == PM_FP64) {
    if (ck == PM_COEFF_FULL) launch_e_t<double, PM_COEFF_FULL>(g, s);
    else launch_e_t<double, PM_COEFF_CB>(g, s);
  } else {
    if (ck == PM_COEFF_FULL) launch_e_t<float, PM_COEFF_FULL>(g, s);
    else if (ck == PM_COEFF_CB) launch_e_t<float, PM_COEFF_CB>(g, s);
    else launch_e_t<float, PM_COEFF_CB_F16>(g, s);
  }
}

}  // namespace pm
