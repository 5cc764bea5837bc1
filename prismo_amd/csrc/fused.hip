// Fused Yee step (H update then E update in one pass) for gfx950.
//
// Same arithmetic as yee.hip (textbook Yee + CPML + ADE), but each step reads
// E(n), H(n-1/2) from one buffer set and writes E(n+1), H(n+1/2) to the other
// (ping-pong), so a step moves every field element through HBM exactly once:
// 12 field accesses + 3 coefficient reads per cell, the algorithmic minimum.
//
// Work decomposition (nx <= 64*VEC, i.e. one x-row per wave64):
//  - a workgroup owns TY x-rows (j0 .. j0+TY-1) for a z-chunk [kb, ke) and
//    has one more wave for the halo row j0+TY, whose H(n+1/2) values the
//    last own row needs for its y-forward differences;
//  - each lane holds VEC consecutive x-cells (vector loads); x-1 / x+1
//    neighbours cross lanes with one shuffle per array;
//  - the z-march keeps planes k-1 and k of E(n) and H(n+1/2) in registers:
//    iteration k computes H(n+1/2) on plane k, then E(n+1) on plane k-1;
//  - LDS carries Ex/Ez of plane k (for the next row's y-backward differences)
//    and Hx/Hz of planes k-1/k (for the previous row's y-forward differences);
//  - the chunk recomputes the H plane ke (not stored) so E(ke-1) needs
//    nothing from the next chunk; H-side CPML memory is ping-ponged too, so
//    recomputed halo values never race with their owner.
// Workgroups map to XCDs in contiguous (y-tile, z-chunk) ranges, so halo
// rows and planes shared by neighbouring tiles are L2 hits.
#include "pm_internal.h"

namespace pm {

template <typename T>
struct Fz {
  // all pointers are the (in, out) pairs of one step
  const T* Ei[3];
  const T* Hi[3];
  T* Eo[3];
  T* Ho[3];
  const T* psh_in[3][2];
  T* psh_out[3][2];
};

__device__ __forceinline__ int slab_of_f(int p, int n, int tlo, int thi) {
  if (p < tlo) return p;
  if (p >= n - thi) return tlo + (p - (n - thi));
  return -1;
}

__device__ __forceinline__ int64_t slab_off_f(const KGrid& g, int a, int s, int i, int j,
                                              int k) {
  if (a == 0) {
    const int S = g.pml[0] + g.pml[1];
    return (int64_t)s + (int64_t)S * ((int64_t)j + (int64_t)g.ny * k);
  } else if (a == 1) {
    const int S = g.pml[2] + g.pml[3];
    return (int64_t)i + (int64_t)g.nx * ((int64_t)s + (int64_t)S * k);
  }
  return (int64_t)i + (int64_t)g.nx * ((int64_t)j + (int64_t)g.ny * s);
}

// H-side CPML term with ping-pong memory; `store` false for halo cells.
template <typename T>
__device__ __forceinline__ T cpml_h(const KGrid& g, const Fz<T>& z, T d, int a, int q, int p,
                                    int s, int i, int j, int k, bool store) {
  const T* b = (const T*)g.pml_bh[a];
  const T* c = (const T*)g.pml_ch[a];
  const T* ik = (const T*)g.pml_ikh[a];
  const int64_t o = slab_off_f(g, a, s, i, j, k);
  const T ps = b[p] * z.psh_in[a][q][o] + c[p] * d;
  if (store) z.psh_out[a][q][o] = ps;
  return d * ik[p] + ps;
}

// E-side CPML term, memory updated in place (E is only computed on own cells).
template <typename T>
__device__ __forceinline__ T cpml_e(const KGrid& g, T d, int a, int q, int p, int s, int i,
                                    int j, int k) {
  const T* b = (const T*)g.pml_be[a];
  const T* c = (const T*)g.pml_ce[a];
  const T* ik = (const T*)g.pml_ike[a];
  T* psi = (T*)g.psi_e[a][q] + slab_off_f(g, a, s, i, j, k);
  const T ps = b[p] * (*psi) + c[p] * d;
  *psi = ps;
  return d * ik[p] + ps;
}

template <typename T>
__device__ __forceinline__ bool ade_f(const KGrid& g, int comp, int i, int j, int k, T curl,
                                      T& e) {
  for (int b = 0; b < g.n_ade; ++b) {
    const KAde& a = g.ade[b];
    if (a.comp != comp || i < a.i0 || i >= a.i1 || j < a.j0 || j >= a.j1 || k < a.k0 ||
        k >= a.k1)
      continue;
    const int64_t l = (int64_t)(i - a.i0) +
                      (int64_t)(a.i1 - a.i0) *
                          ((int64_t)(j - a.j0) + (int64_t)(a.j1 - a.j0) * (k - a.k0));
    const T s = (T)a.cb_scale * curl;
    const T einf = (T)a.eps_inf;
    T* P = (T*)a.p;
    if (a.kind == 1) {
      T* Pm = (T*)a.p_prev;
      const T p0 = P[l], pm = Pm[l];
      const T en = ((einf - (T)a.c1) * e + ((T)1 - (T)a.c2) * p0 - (T)a.c3 * pm + s) /
                   (einf + (T)a.c0);
      P[l] = (T)a.c0 * en + (T)a.c1 * e + (T)a.c2 * p0 + (T)a.c3 * pm;
      Pm[l] = p0;
      e = en;
    } else {
      const T J = P[l];
      const T hdt = (T)(0.5 * g.dt);
      const T en = (einf * e + s - hdt * ((T)1 + (T)a.c1) * J) / (einf + hdt * (T)a.c0);
      P[l] = (T)a.c0 * en + (T)a.c1 * J;
      e = en;
    }
    return true;
  }
  return false;
}

template <typename T, int VEC>
struct Vec {
  using type = T __attribute__((ext_vector_type(VEC)));
};

template <typename T, int VEC>
__device__ __forceinline__ void vld(T (&d)[VEC], const T* p) {
  using VT = typename Vec<T, VEC>::type;
  const VT v = *(const VT*)p;
#pragma unroll
  for (int e = 0; e < VEC; ++e) d[e] = v[e];
}

template <typename T, int VEC>
__device__ __forceinline__ void vst(T* p, const T (&s)[VEC]) {
  using VT = typename Vec<T, VEC>::type;
  VT v;
#pragma unroll
  for (int e = 0; e < VEC; ++e) v[e] = s[e];
  *(VT*)p = v;
}

template <typename T, int VEC>
__device__ __forceinline__ void vzero(T (&d)[VEC]) {
#pragma unroll
  for (int e = 0; e < VEC; ++e) d[e] = (T)0;
}

template <typename T, int CK, int VEC>
__device__ __forceinline__ void load_cb_vec(const KGrid& g, int c, int64_t off, T (&cb)[VEC]) {
  if constexpr (CK == PM_COEFF_CB_F16) {
    using HV = _Float16 __attribute__((ext_vector_type(VEC)));
    const HV h = *(const HV*)((const _Float16*)g.cb[c] + off);
    const T s = (T)g.cb_scale;
#pragma unroll
    for (int e = 0; e < VEC; ++e) cb[e] = s * (T)h[e];
  } else {
    vld<T, VEC>(cb, (const T*)g.cb[c] + off);
  }
}

template <typename T>
__device__ __forceinline__ T shfl_up1(T v) {
  return __shfl_up(v, 1, 64);
}
template <typename T>
__device__ __forceinline__ T shfl_down1(T v) {
  return __shfl_down(v, 1, 64);
}

template <typename T, int CK, int VEC, int TY>
__global__ __launch_bounds__(64 * (TY + 1)) void k_yee_fused(const KGrid g, const Fz<T> z,
                                                             int kchunk, int ntile_y,
                                                             int nblocks) {
  constexpr int RW = 64 * VEC;  // cells per row held by one wave
  __shared__ T lds_e[TY + 1][2][RW];     // Ex, Ez of plane k
  __shared__ T lds_h[2][TY + 1][2][RW];  // Hx, Hz of planes (k & 1)

  // XCD-contiguous tile order: hardware id b runs on XCD b % 8
  const int b = blockIdx.x;
  const int per = (nblocks + 7) / 8;
  const int L = (b % 8) * per + b / 8;
  if (L >= nblocks) return;  // whole workgroup exits together
  const int ty = L % ntile_y;
  const int kc = L / ntile_y;

  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;  // row within the tile, TY = halo row
  const int j = ty * TY + w;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const bool row_ok = j < ny;
  const bool own = (w < TY) && row_ok;
  const int kb = kc * kchunk;
  const int ke = min(kb + kchunk, nz);
  const int kh = min(ke, nz - 1);  // last H plane computed (ke itself is a halo)
  const int i0 = lane * VEC;
  const T idx = (T)g.idx, idy = (T)g.idy, idz = (T)g.idz;
  const int64_t sy = g.sy, sz = g.sz;
  const int64_t rowoff = (int64_t)j * sy + i0;

  const int sys = g.has_pml ? slab_of_f(j, ny, g.pml[2], g.pml[3]) : -1;
  int sxs[VEC];
#pragma unroll
  for (int e = 0; e < VEC; ++e)
    sxs[e] = (g.has_pml && i0 + e < nx) ? slab_of_f(i0 + e, nx, g.pml[0], g.pml[1]) : -1;
  bool lane_pml_x = false;
#pragma unroll
  for (int e = 0; e < VEC; ++e) lane_pml_x |= sxs[e] >= 0;

  // E(n) planes k-1 (m) and k; H(n+1/2) planes k-1 (p) and k
  T ex[VEC], ey[VEC], ez[VEC], exm[VEC], eym[VEC], ezm[VEC];
  T hx[VEC], hy[VEC], hz[VEC], hxp[VEC], hyp[VEC], hzp[VEC];
  vzero(exm);
  vzero(eym);
  vzero(ezm);
  vzero(hxp);
  vzero(hyp);
  vzero(hzp);
  if (row_ok && kb > 0) {
    const int64_t o = rowoff + (int64_t)(kb - 1) * sz;
    vld(exm, z.Ei[0] + o);
    vld(eym, z.Ei[1] + o);
  }

  for (int k = kb; k <= kh; ++k) {
    const int64_t off = rowoff + (int64_t)k * sz;
    // (a) E(n) on plane k
    if (row_ok) {
      vld(ex, z.Ei[0] + off);
      vld(ey, z.Ei[1] + off);
      vld(ez, z.Ei[2] + off);
    } else {
      vzero(ex);
      vzero(ey);
      vzero(ez);
    }
    // (b) share Ex, Ez of this row
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      lds_e[w][0][i0 + e] = ex[e];
      lds_e[w][1][i0 + e] = ez[e];
    }
    __syncthreads();
    // (c) row j-1
    T exj[VEC], ezj[VEC];
    if (w > 0) {
#pragma unroll
      for (int e = 0; e < VEC; ++e) {
        exj[e] = lds_e[w - 1][0][i0 + e];
        ezj[e] = lds_e[w - 1][1][i0 + e];
      }
    } else if (row_ok && j > 0) {
      vld(exj, z.Ei[0] + off - sy);
      vld(ezj, z.Ei[2] + off - sy);
    } else {
      vzero(exj);
      vzero(ezj);
    }
    // (d) column i-1
    T ezi[VEC], eyi[VEC];
    {
      const T ez_l = shfl_up1(ez[VEC - 1]);
      const T ey_l = shfl_up1(ey[VEC - 1]);
      ezi[0] = lane > 0 ? ez_l : (T)0;
      eyi[0] = lane > 0 ? ey_l : (T)0;
#pragma unroll
      for (int e = 1; e < VEC; ++e) {
        ezi[e] = ez[e - 1];
        eyi[e] = ey[e - 1];
      }
    }
    // (f) H(n+1/2) on plane k
    const bool store_h = own && k < ke;
    const int szs = g.has_pml ? slab_of_f(k, nz, g.pml[4], g.pml[5]) : -1;
    if (row_ok) {
      T hxi[VEC], hyi[VEC], hzi[VEC];
      vld(hxi, z.Hi[0] + off);
      vld(hyi, z.Hi[1] + off);
      vld(hzi, z.Hi[2] + off);
      const bool pml_any = lane_pml_x || sys >= 0 || szs >= 0;
      const T db = (T)g.db_const;
#pragma unroll
      for (int e = 0; e < VEC; ++e) {
        const int i = i0 + e;
        T c1 = (ez[e] - ezj[e]) * idy;  // Hx: dEz/dy - dEy/dz
        T c2 = (ey[e] - eym[e]) * idz;
        T d1 = (ex[e] - exm[e]) * idz;  // Hy: dEx/dz - dEz/dx
        T d2 = (ez[e] - ezi[e]) * idx;
        T f1 = (ey[e] - eyi[e]) * idx;  // Hz: dEy/dx - dEx/dy
        T f2 = (ex[e] - exj[e]) * idy;
        if (pml_any && i < nx) {  // padding lanes never touch the slab arrays
          const bool st = store_h;
          if (sys >= 0) {
            c1 = cpml_h(g, z, c1, 1, 1, j, sys, i, j, k, st && i < nx - 1);
            f2 = cpml_h(g, z, f2, 1, 0, j, sys, i, j, k, st && k < nz - 1);
          }
          if (szs >= 0) {
            c2 = cpml_h(g, z, c2, 2, 0, k, szs, i, j, k, st && i < nx - 1);
            d1 = cpml_h(g, z, d1, 2, 1, k, szs, i, j, k, st && j < ny - 1);
          }
          if (sxs[e] >= 0) {
            d2 = cpml_h(g, z, d2, 0, 0, i, sxs[e], i, j, k, st && j < ny - 1);
            f1 = cpml_h(g, z, f1, 0, 1, i, sxs[e], i, j, k, st && k < nz - 1);
          }
        }
        if constexpr (CK == PM_COEFF_FULL) {
          const T dax = ((const T*)g.da[0])[off + e], dbx = ((const T*)g.db[0])[off + e];
          const T day = ((const T*)g.da[1])[off + e], dby = ((const T*)g.db[1])[off + e];
          const T daz = ((const T*)g.da[2])[off + e], dbz = ((const T*)g.db[2])[off + e];
          hx[e] = dax * hxi[e] - dbx * (c1 - c2);
          hy[e] = day * hyi[e] - dby * (d1 - d2);
          hz[e] = daz * hzi[e] - dbz * (f1 - f2);
        } else {
          hx[e] = hxi[e] - db * (c1 - c2);
          hy[e] = hyi[e] - db * (d1 - d2);
          hz[e] = hzi[e] - db * (f1 - f2);
        }
        // outside a component's extent the buffer holds zeros
        if (i >= nx - 1) hx[e] = 0;
        if (i >= nx || j >= ny - 1) hy[e] = 0;
        if (i >= nx || k >= nz - 1) hz[e] = 0;
      }
      if (store_h) {
        vst(z.Ho[0] + off, hx);
        vst(z.Ho[1] + off, hy);
        vst(z.Ho[2] + off, hz);
      }
    } else {
      vzero(hx);
      vzero(hy);
      vzero(hz);
    }
    // (g) share Hx, Hz of plane k
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      lds_h[k & 1][w][0][i0 + e] = hx[e];
      lds_h[k & 1][w][1][i0 + e] = hz[e];
    }
    __syncthreads();
    // (h) E(n+1) on plane k-1
    if (k > kb && own) {
      const int kk = k - 1;
      const int64_t o = rowoff + (int64_t)kk * sz;
      const int szz = g.has_pml ? slab_of_f(kk, nz, g.pml[4], g.pml[5]) : -1;
      T hyx[VEC], hzx[VEC], hxy[VEC], hzy[VEC];
      {
        const T hy_r = shfl_down1(hyp[0]);
        const T hz_r = shfl_down1(hzp[0]);
#pragma unroll
        for (int e = 0; e < VEC - 1; ++e) {
          hyx[e] = hyp[e + 1];
          hzx[e] = hzp[e + 1];
        }
        hyx[VEC - 1] = hy_r;
        hzx[VEC - 1] = hz_r;
      }
#pragma unroll
      for (int e = 0; e < VEC; ++e) {
        hxy[e] = lds_h[kk & 1][w + 1][0][i0 + e];
        hzy[e] = lds_h[kk & 1][w + 1][1][i0 + e];
      }
      T cbx[VEC], cby[VEC], cbz[VEC];
      load_cb_vec<T, CK, VEC>(g, 0, o, cbx);
      load_cb_vec<T, CK, VEC>(g, 1, o, cby);
      load_cb_vec<T, CK, VEC>(g, 2, o, cbz);
      T nex[VEC], ney[VEC], nez[VEC];
      const bool pml_any = lane_pml_x || sys >= 0 || szz >= 0;
#pragma unroll
      for (int e = 0; e < VEC; ++e) {
        const int i = i0 + e;
        T a1 = (hzy[e] - hzp[e]) * idy;  // Ex: dHz/dy - dHy/dz
        T a2 = (hy[e] - hyp[e]) * idz;
        T b1 = (hx[e] - hxp[e]) * idz;   // Ey: dHx/dz - dHz/dx
        T b2 = (hzx[e] - hzp[e]) * idx;
        T c1 = (hyx[e] - hyp[e]) * idx;  // Ez: dHy/dx - dHx/dy
        T c2 = (hxy[e] - hxp[e]) * idy;
        const bool exok = i < nx && j < ny - 1 && kk < nz - 1;
        const bool eyok = i < nx - 1 && kk < nz - 1;
        const bool ezok = i < nx - 1 && j < ny - 1;
        if (pml_any) {
          if (sys >= 0) {
            if (exok) a1 = cpml_e(g, a1, 1, 1, j, sys, i, j, kk);
            if (ezok) c2 = cpml_e(g, c2, 1, 0, j, sys, i, j, kk);
          }
          if (szz >= 0) {
            if (exok) a2 = cpml_e(g, a2, 2, 0, kk, szz, i, j, kk);
            if (eyok) b1 = cpml_e(g, b1, 2, 1, kk, szz, i, j, kk);
          }
          if (sxs[e] >= 0) {
            if (eyok) b2 = cpml_e(g, b2, 0, 0, i, sxs[e], i, j, kk);
            if (ezok) c1 = cpml_e(g, c1, 0, 1, i, sxs[e], i, j, kk);
          }
        }
        T vx = exm[e], vy = eym[e], vz = ezm[e];
        const T cx = a1 - a2, cy = b1 - b2, cz = c1 - c2;
        if (!(g.n_ade && ade_f<T>(g, 0, i, j, kk, cx, vx))) {
          if constexpr (CK == PM_COEFF_FULL) vx = ((const T*)g.ca[0])[o + e] * vx + cbx[e] * cx;
          else vx = vx + cbx[e] * cx;
        }
        if (!(g.n_ade && ade_f<T>(g, 1, i, j, kk, cy, vy))) {
          if constexpr (CK == PM_COEFF_FULL) vy = ((const T*)g.ca[1])[o + e] * vy + cby[e] * cy;
          else vy = vy + cby[e] * cy;
        }
        if (!(g.n_ade && ade_f<T>(g, 2, i, j, kk, cz, vz))) {
          if constexpr (CK == PM_COEFF_FULL) vz = ((const T*)g.ca[2])[o + e] * vz + cbz[e] * cz;
          else vz = vz + cbz[e] * cz;
        }
        nex[e] = exok ? vx : (T)0;
        ney[e] = eyok ? vy : (T)0;
        nez[e] = ezok ? vz : (T)0;
      }
      vst(z.Eo[0] + o, nex);
      vst(z.Eo[1] + o, ney);
      vst(z.Eo[2] + o, nez);
    }
    // rotate planes
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      exm[e] = ex[e];
      eym[e] = ey[e];
      ezm[e] = ez[e];
      hxp[e] = hx[e];
      hyp[e] = hy[e];
      hzp[e] = hz[e];
    }
  }
  // the last own plane when the chunk ends at the grid top: only Ez exists
  // there (Ex, Ey end at nz-2) and it has no z differences
  if (ke == nz && own) {
    const int kk = nz - 1;
    const int64_t o = rowoff + (int64_t)kk * sz;
    T hyx[VEC], hxy[VEC];
    const T hy_r = shfl_down1(hyp[0]);
#pragma unroll
    for (int e = 0; e < VEC - 1; ++e) hyx[e] = hyp[e + 1];
    hyx[VEC - 1] = hy_r;
#pragma unroll
    for (int e = 0; e < VEC; ++e) hxy[e] = lds_h[kk & 1][w + 1][0][i0 + e];
    T cbz[VEC];
    load_cb_vec<T, CK, VEC>(g, 2, o, cbz);
    T nez[VEC], zero[VEC];
    vzero(zero);
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      const int i = i0 + e;
      T c1 = (hyx[e] - hyp[e]) * idx;
      T c2 = (hxy[e] - hxp[e]) * idy;
      const bool ezok = i < nx - 1 && j < ny - 1;
      if (ezok && g.has_pml) {
        if (sys >= 0) c2 = cpml_e(g, c2, 1, 0, j, sys, i, j, kk);
        if (sxs[e] >= 0) c1 = cpml_e(g, c1, 0, 1, i, sxs[e], i, j, kk);
      }
      T vz = ezm[e];
      const T cz = c1 - c2;
      if (!(g.n_ade && ade_f<T>(g, 2, i, j, kk, cz, vz))) {
        if constexpr (CK == PM_COEFF_FULL) vz = ((const T*)g.ca[2])[o + e] * vz + cbz[e] * cz;
        else vz = vz + cbz[e] * cz;
      }
      nez[e] = ezok ? vz : (T)0;
    }
    vst(z.Eo[0] + o, zero);
    vst(z.Eo[1] + o, zero);
    vst(z.Eo[2] + o, nez);
  }
}

// ---------------------------------------------------------------- launch
template <typename T, int CK, int VEC, int TY>
static void launch_fused_t(const KGrid& g, const Fz<T>& z, int kchunk, hipStream_t s) {
  const int ntile_y = (g.ny + TY - 1) / TY;
  const int nchunk = (g.nz + kchunk - 1) / kchunk;
  const int nb = ntile_y * nchunk;
  const int nlaunch = ((nb + 7) / 8) * 8;
  hipLaunchKernelGGL((k_yee_fused<T, CK, VEC, TY>), dim3(nlaunch), dim3(64 * (TY + 1)), 0, s, g,
                     z, kchunk, ntile_y, nb);
}

template <typename T, int VEC, int TY>
constexpr bool fused_fits() {
  // static LDS: Ex/Ez of one plane + Hx/Hz of two planes, TY + 1 rows
  return (size_t)3 * 2 * (TY + 1) * 64 * VEC * sizeof(T) <= 65536;
}

template <typename T, int CK, int VEC, int TY>
static int launch_if_fits(const KGrid& g, const Fz<T>& z, int kchunk, hipStream_t s) {
  if constexpr (fused_fits<T, VEC, TY>()) {
    launch_fused_t<T, CK, VEC, TY>(g, z, kchunk, s);
    return 0;
  } else {
    return -1;
  }
}

template <typename T, int CK, int VEC>
static int launch_fused_v(const KGrid& g, const Fz<T>& z, int ty, int kchunk, hipStream_t s) {
  if (ty == 2) return launch_if_fits<T, CK, VEC, 2>(g, z, kchunk, s);
  if (ty == 4) return launch_if_fits<T, CK, VEC, 4>(g, z, kchunk, s);
  if (ty == 8) return launch_if_fits<T, CK, VEC, 8>(g, z, kchunk, s);
  return -1;
}

template <typename T, int CK>
static int launch_fused_ck(const KGrid& g, const Fz<T>& z, int vec, int ty, int kchunk,
                           hipStream_t s) {
  if (vec == 2) return launch_fused_v<T, CK, 2>(g, z, ty, kchunk, s);
  if (vec == 4) return launch_fused_v<T, CK, 4>(g, z, ty, kchunk, s);
  if (vec == 8) return launch_fused_v<T, CK, 8>(g, z, ty, kchunk, s);
  return -1;
}

bool fused_supported(int precision, int vec, int ty) {
  const size_t t = precision == PM_FP64 ? 8 : 4;
  if ((vec != 2 && vec != 4 && vec != 8) || (ty != 2 && ty != 4 && ty != 8)) return false;
  return (size_t)3 * 2 * (ty + 1) * 64 * vec * t <= 65536;
}

// Returns 0 on launch, -1 when the (vec, ty) pair is not instantiated.
int launch_yee_fused(const KGrid& g, int precision, int ck, const FusedBufs& fb, int vec, int ty,
                     int kchunk, hipStream_t s) {
  if (precision == PM_FP32) {
    Fz<float> z{};
    for (int a = 0; a < 3; ++a) {
      z.Ei[a] = (const float*)fb.Ei[a];
      z.Hi[a] = (const float*)fb.Hi[a];
      z.Eo[a] = (float*)fb.Eo[a];
      z.Ho[a] = (float*)fb.Ho[a];
      for (int q = 0; q < 2; ++q) {
        z.psh_in[a][q] = (const float*)fb.psh_in[a][q];
        z.psh_out[a][q] = (float*)fb.psh_out[a][q];
      }
    }
    if (ck == PM_COEFF_FULL) return launch_fused_ck<float, PM_COEFF_FULL>(g, z, vec, ty, kchunk, s);
    if (ck == PM_COEFF_CB) return launch_fused_ck<float, PM_COEFF_CB>(g, z, vec, ty, kchunk, s);
    return launch_fused_ck<float, PM_COEFF_CB_F16>(g, z, vec, ty, kchunk, s);
  }
  Fz<double> z{};
  for (int a = 0; a < 3; ++a) {
    z.Ei[a] = (const double*)fb.Ei[a];
    z.Hi[a] = (const double*)fb.Hi[a];
    z.Eo[a] = (double*)fb.Eo[a];
    z.Ho[a] = (double*)fb.Ho[a];
    for (int q = 0; q < 2; ++q) {
      z.psh_in[a][q] = (const double*)fb.psh_in[a][q];
      z.psh_out[a][q] = (double*)fb.psh_out[a][q];
    }
  }
  if (ck == PM_COEFF_FULL) return launch_fused_ck<double, PM_COEFF_FULL>(g, z, vec, ty, kchunk, s);
  return launch_fused_ck<double, PM_COEFF_CB>(g, z, vec, ty, kchunk, s);
}

}  // namespace pm
