// Fused Yee step (H update then E update in one pass) for gfx950.
//
// Same arithmetic as yee.hip (textbook Yee + CPML + ADE), but each step reads
// E(n), H(n-1/2) from one buffer set and writes E(n+1), H(n+1/2) to the other
// (ping-pong), so a step moves every field element through HBM once: 12 field
// accesses + 3 coefficient reads per cell, the algorithmic minimum.
//
// Decomposition (one padded x-row of RW = 64*VEC*W cells, RW >= nx):
//  - a workgroup owns TY rows (j0 .. j0+TY-1) of a z-chunk [kb, ke); W waves
//    per row, VEC consecutive x-cells per lane (vector loads), plus W waves
//    for the halo row j0+TY whose H(n+1/2) the last own row needs;
//  - the z-march keeps E(n) and H(n+1/2) of planes k-1 and k in registers:
//    iteration k computes H(n+1/2) on plane k, then E(n+1) on plane k-1, and
//    prefetches plane k+1 from HBM while plane k is in flight;
//  - all in-plane neighbours come from LDS: E(n) of plane k for the i-1/j-1
//    differences of H, H(n+1/2) of planes k-1/k (double buffer) for the
//    i+1/j+1 differences of E;
//  - the chunk recomputes H on plane ke (not stored) so E(ke-1) needs nothing
//    from the next chunk; H-side CPML memory is ping-ponged as well, so the
//    recomputed halo values never race with their owner;
//  - workgroups map to XCDs in contiguous ranges of (y-tile, z-chunk) in the
//    order picked by the tuner, so shared halo rows / planes are L2 hits.
#pragma once
#include "pm_internal.h"

namespace pm {

template <typename T>
struct Fz {
  const T* Ei[3];
  const T* Hi[3];
  T* Eo[3];
  T* Ho[3];
  const T* psh_in[3][2];
  T* psh_out[3][2];
};

__device__ __forceinline__ int fslab(int p, int n, int tlo, int thi) {
  if (p < tlo) return p;
  if (p >= n - thi) return tlo + (p - (n - thi));
  return -1;
}

__device__ __forceinline__ int64_t fslab_off(const KGrid& g, int a, int s, int i, int j, int k) {
  if (a == 0) {
    const int S = g.pml[0] + g.pml[1];
    return (int64_t)s + (int64_t)S * ((int64_t)j + (int64_t)g.ny * k);
  } else if (a == 1) {
    const int S = g.pml[2] + g.pml[3];
    return (int64_t)i + (int64_t)g.nx * ((int64_t)s + (int64_t)S * k);
  }
  return (int64_t)i + (int64_t)g.nx * ((int64_t)j + (int64_t)g.ny * s);
}

// H-side CPML term, ping-pong memory; `store` false for halo cells.
template <typename T>
__device__ __forceinline__ T fcpml_h(const KGrid& g, const Fz<T>& z, T d, int a, int q, int p,
                                     int s, int i, int j, int k, bool store) {
  const T* b = (const T*)g.pml_bh[a];
  const T* c = (const T*)g.pml_ch[a];
  const T* ik = (const T*)g.pml_ikh[a];
  const int64_t o = fslab_off(g, a, s, i, j, k);
  const T ps = b[p] * z.psh_in[a][q][o] + c[p] * d;
  if (store) z.psh_out[a][q][o] = ps;
  return d * ik[p] + ps;
}

// As fcpml_h with the memory value already loaded (offset o precomputed).
template <typename T>
__device__ __forceinline__ T fcpml_hq(const KGrid& g, const Fz<T>& z, T d, T psi_old, int a,
                                      int q, int p, int64_t o, bool store) {
  const T ps = ((const T*)g.pml_bh[a])[p] * psi_old + ((const T*)g.pml_ch[a])[p] * d;
  if (store) z.psh_out[a][q][o] = ps;
  return d * ((const T*)g.pml_ikh[a])[p] + ps;
}

// E-side term with the memory value already loaded; stores in place.
template <typename T>
__device__ __forceinline__ T fcpml_eq(const KGrid& g, T d, T psi_old, int a, int q, int p,
                                      int64_t o) {
  const T ps = ((const T*)g.pml_be[a])[p] * psi_old + ((const T*)g.pml_ce[a])[p] * d;
  ((T*)g.psi_e[a][q])[o] = ps;
  return d * ((const T*)g.pml_ike[a])[p] + ps;
}

// E-side CPML term, in place (E is computed on own cells only).
template <typename T>
__device__ __forceinline__ T fcpml_e(const KGrid& g, T d, int a, int q, int p, int s, int i,
                                     int j, int k) {
  const T* b = (const T*)g.pml_be[a];
  const T* c = (const T*)g.pml_ce[a];
  const T* ik = (const T*)g.pml_ike[a];
  T* psi = (T*)g.psi_e[a][q] + fslab_off(g, a, s, i, j, k);
  const T ps = b[p] * (*psi) + c[p] * d;
  *psi = ps;
  return d * ik[p] + ps;
}

// Dispersive E update from the device box table (see yee.hip for the math).
template <typename T>
struct AdeOut {
  T e;
  bool hit;
};

template <typename T>
__device__ __noinline__ AdeOut<T> fade(const KAde* tab, int n_ade, int comp, int i, int j,
                                       int k, T curl, T e) {
#pragma unroll 1
  for (int b = 0; b < n_ade; ++b) {
    const KAde& a = tab[b];
    if (a.comp != comp || i < a.i0 || i >= a.i1 || j < a.j0 || j >= a.j1 || k < a.k0 ||
        k >= a.k1)
      continue;
    const int64_t l = (int64_t)(i - a.i0) +
                      (int64_t)(a.i1 - a.i0) *
                          ((int64_t)(j - a.j0) + (int64_t)(a.j1 - a.j0) * (k - a.k0));
    return AdeOut<T>{ade_cell<T>(a, l, curl, e), true};
  }
  return AdeOut<T>{e, false};
}

// Does row j of plane k cut any dispersive box? (same for every lane)
__device__ __forceinline__ bool row_has_ade(const KGrid& g, int j, int k) {
  for (int b = 0; b < g.n_ade; ++b) {
    const KAde& a = g.ade[b];
    if (j >= a.j0 && j < a.j1 && k >= a.k0 && k < a.k1) return true;
  }
  return false;
}

template <typename T, int VEC>
struct FVec {
  using type = T __attribute__((ext_vector_type(VEC)));
};

template <typename T, int VEC>
__device__ __forceinline__ void fld(T (&d)[VEC], const T* p) {
  if constexpr (VEC == 1) {
    d[0] = *p;
  } else {
    using VT = typename FVec<T, VEC>::type;
    const VT v = *(const VT*)p;
#pragma unroll
    for (int e = 0; e < VEC; ++e) d[e] = v[e];
  }
}

template <typename T, int VEC>
__device__ __forceinline__ void fst(T* p, const T (&s)[VEC]) {
  if constexpr (VEC == 1) {
    *p = s[0];
  } else {
    using VT = typename FVec<T, VEC>::type;
    VT v;
#pragma unroll
    for (int e = 0; e < VEC; ++e) v[e] = s[e];
    *(VT*)p = v;
  }
}

template <typename T, int VEC>
__device__ __forceinline__ void fzero(T (&d)[VEC]) {
#pragma unroll
  for (int e = 0; e < VEC; ++e) d[e] = (T)0;
}

template <typename T, int CK, int VEC>
__device__ __forceinline__ void fcb(const KGrid& g, int c, uint32_t off, T (&cb)[VEC]) {
  if constexpr (CK == PM_COEFF_CB_F16) {
    const _Float16* h = (const _Float16*)g.cb[c] + off;
    const T s = (T)g.cb_scale;
    if constexpr (VEC == 1) {
      cb[0] = s * (T)h[0];
    } else {
      using HV = _Float16 __attribute__((ext_vector_type(VEC)));
      const HV v = *(const HV*)h;
#pragma unroll
      for (int e = 0; e < VEC; ++e) cb[e] = s * (T)v[e];
    }
  } else {
    fld<T, VEC>(cb, (const T*)g.cb[c] + off);
  }
}

template <typename T, int CK, int VEC, int W, int TY>
__global__ __launch_bounds__(64 * W * (TY + 1)) void k_yee_fused(const KGrid g, const Fz<T> z,
                                                                 int kchunk, int ntile_y,
                                                                 int nchunk, int nblocks,
                                                                 int order) {
  constexpr int RW = 64 * VEC * W;
  __shared__ T le[3][TY + 1][RW];     // E(n) of plane k
  __shared__ T lh[2][3][TY + 1][RW];  // H(n+1/2) of planes with parity 0 / 1

  const int b = blockIdx.x;
  const int per = (nblocks + 7) / 8;
  // orders 0/1: XCD b % 8 runs a contiguous range of tiles (halo reuse in its
  // L2); order 2: hardware round-robin (slab tiles spread over all XCDs)
  const int L = order == 2 ? b : (b % 8) * per + b / 8;
  if (L >= nblocks) return;
  int ty, kc;
  if (order != 1) {
    ty = L % ntile_y;
    kc = L / ntile_y;
  } else {
    kc = L % nchunk;
    ty = L / nchunk;
  }

  const int tid = threadIdx.x;
  // the row index is the same for a whole wave: make it scalar so row-wise
  // quantities (y profiles, slab indices) live in SGPRs and load as scalars
  const int w = __builtin_amdgcn_readfirstlane(tid / (64 * W));  // TY = halo row
  const int x0 = (tid % (64 * W)) * VEC;
  const int j = ty * TY + w;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const bool row_ok = j < ny;
  const bool own = (w < TY) && row_ok;
  const int kb = kc * kchunk;
  const int ke = min(kb + kchunk, nz);
  const int kh = min(ke, nz - 1);
  const T idx = (T)g.idx, idy = (T)g.idy, idz = (T)g.idz;
  // 32-bit element offsets (the host keeps every field array below 4 GB)
  const uint32_t sy = (uint32_t)g.sy, sz = (uint32_t)g.sz;
  const uint32_t rowoff = (uint32_t)j * sy + (uint32_t)x0;
  const bool first_row = (w == 0) && row_ok && j > 0;
  const bool pml = g.has_pml;

  // ---- CPML profiles: x per lane (registers), y per row (scalars) -------
  const int sys = (pml && row_ok) ? fslab(j, ny, g.pml[2], g.pml[3]) : -1;
  int sxs[VEC];
  bool lane_pml_x = false;
  T xbh[VEC], xch[VEC], xikh[VEC], xbe[VEC], xce[VEC], xike[VEC];
#pragma unroll
  for (int e = 0; e < VEC; ++e) {
    const int i = x0 + e;
    sxs[e] = (pml && i < nx) ? fslab(i, nx, g.pml[0], g.pml[1]) : -1;
    lane_pml_x |= sxs[e] >= 0;
    xbh[e] = xbe[e] = (T)1;
    xch[e] = xce[e] = (T)0;
    xikh[e] = xike[e] = (T)1;
    if (sxs[e] >= 0) {
      xbh[e] = ((const T*)g.pml_bh[0])[i];
      xch[e] = ((const T*)g.pml_ch[0])[i];
      xikh[e] = ((const T*)g.pml_ikh[0])[i];
      xbe[e] = ((const T*)g.pml_be[0])[i];
      xce[e] = ((const T*)g.pml_ce[0])[i];
      xike[e] = ((const T*)g.pml_ike[0])[i];
    }
  }
  T ybh = 1, ych = 0, yikh = 1, ybe = 1, yce = 0, yike = 1;
  if (sys >= 0) {
    ybh = ((const T*)g.pml_bh[1])[j];
    ych = ((const T*)g.pml_ch[1])[j];
    yikh = ((const T*)g.pml_ikh[1])[j];
    ybe = ((const T*)g.pml_be[1])[j];
    yce = ((const T*)g.pml_ce[1])[j];
    yike = ((const T*)g.pml_ike[1])[j];
  }
  const T* __restrict__ zbh = (const T*)g.pml_bh[2];
  const T* __restrict__ zch = (const T*)g.pml_ch[2];
  const T* __restrict__ zikh = (const T*)g.pml_ikh[2];
  const T* __restrict__ zbe = (const T*)g.pml_be[2];
  const T* __restrict__ zce = (const T*)g.pml_ce[2];
  const T* __restrict__ zike = (const T*)g.pml_ike[2];

  // H-side CPML memory of one plane: [x0 x1 y0 y1 z0 z1] per cell
  auto load_qh = [&](int k, T (&q)[6][VEC]) {
    const int sk = (pml && row_ok) ? fslab(k, nz, g.pml[4], g.pml[5]) : -1;
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      const int i = x0 + e;
#pragma unroll
      for (int t = 0; t < 6; ++t) q[t][e] = (T)0;
      if (sxs[e] >= 0) {
        const int64_t o = fslab_off(g, 0, sxs[e], i, j, k);
        q[0][e] = z.psh_in[0][0][o];
        q[1][e] = z.psh_in[0][1][o];
      }
      if (sys >= 0 && i < nx) {
        const int64_t o = fslab_off(g, 1, sys, i, j, k);
        q[2][e] = z.psh_in[1][0][o];
        q[3][e] = z.psh_in[1][1][o];
      }
      if (sk >= 0 && i < nx) {
        const int64_t o = fslab_off(g, 2, sk, i, j, k);
        q[4][e] = z.psh_in[2][0][o];
        q[5][e] = z.psh_in[2][1][o];
      }
    }
  };

  T ex[VEC], ey[VEC], ez[VEC], exm[VEC], eym[VEC], ezm[VEC];
  T hx[VEC], hy[VEC], hz[VEC], hxp[VEC], hyp[VEC], hzp[VEC];
  T hxi[VEC], hyi[VEC], hzi[VEC];
  T exu[VEC], ezu[VEC];  // row j-1 of plane k (first row only)
  T qh[6][VEC];
  fzero(exm);
  fzero(eym);
  fzero(ezm);
  fzero(hxp);
  fzero(hyp);
  fzero(hzp);
  fzero(exu);
  fzero(ezu);
  if (row_ok && kb > 0) {
    const uint32_t o = rowoff + (uint32_t)(kb - 1) * sz;
    fld(exm, z.Ei[0] + o);
    fld(eym, z.Ei[1] + o);
  }
  if (row_ok) {
    const uint32_t o = rowoff + (uint32_t)kb * sz;
    fld(ex, z.Ei[0] + o);
    fld(ey, z.Ei[1] + o);
    fld(ez, z.Ei[2] + o);
    fld(hxi, z.Hi[0] + o);
    fld(hyi, z.Hi[1] + o);
    fld(hzi, z.Hi[2] + o);
    if (first_row) {
      fld(exu, z.Ei[0] + o - sy);
      fld(ezu, z.Ei[2] + o - sy);
    }
  } else {
    fzero(ex);
    fzero(ey);
    fzero(ez);
    fzero(hxi);
    fzero(hyi);
    fzero(hzi);
  }
  if (pml) load_qh(kb, qh);

  for (int k = kb; k <= kh; ++k) {
    const uint32_t off = rowoff + (uint32_t)k * sz;
    const bool do_e = k > kb && own;
    const int kk = k - 1;
    // (1) what this iteration's E update needs, issued ahead of the next
    //     plane's prefetch: vector memory counters retire in order, so a
    //     later wait on these does not wait on the prefetch
    T cbx[VEC], cby[VEC], cbz[VEC];
    T qe[6][VEC];
    const int szz = (pml && do_e) ? fslab(kk, nz, g.pml[4], g.pml[5]) : -1;
    if (do_e) {
      fcb<T, CK, VEC>(g, 0, off - sz, cbx);
      fcb<T, CK, VEC>(g, 1, off - sz, cby);
      fcb<T, CK, VEC>(g, 2, off - sz, cbz);
    }
    if (pml) {
#pragma unroll
      for (int e = 0; e < VEC; ++e) {
        const int i = x0 + e;
#pragma unroll
        for (int t = 0; t < 6; ++t) qe[t][e] = (T)0;
        if (!do_e || i >= nx) continue;
        if (sxs[e] >= 0) {
          const int64_t o = fslab_off(g, 0, sxs[e], i, j, kk);
          qe[0][e] = ((const T*)g.psi_e[0][0])[o];
          qe[1][e] = ((const T*)g.psi_e[0][1])[o];
        }
        if (sys >= 0) {
          const int64_t o = fslab_off(g, 1, sys, i, j, kk);
          qe[2][e] = ((const T*)g.psi_e[1][0])[o];
          qe[3][e] = ((const T*)g.psi_e[1][1])[o];
        }
        if (szz >= 0) {
          const int64_t o = fslab_off(g, 2, szz, i, j, kk);
          qe[4][e] = ((const T*)g.psi_e[2][0])[o];
          qe[5][e] = ((const T*)g.psi_e[2][1])[o];
        }
      }
    }
    // (2) prefetch plane k+1: fields, row j-1, H-side CPML memory
    T nex[VEC], ney[VEC], nez[VEC], nhx[VEC], nhy[VEC], nhz[VEC], nexu[VEC], nezu[VEC];
    T nqh[6][VEC];
    const bool pf = row_ok && k + 1 <= kh;
    if (pf) {
      const uint32_t o = off + sz;
      fld(nex, z.Ei[0] + o);
      fld(ney, z.Ei[1] + o);
      fld(nez, z.Ei[2] + o);
      fld(nhx, z.Hi[0] + o);
      fld(nhy, z.Hi[1] + o);
      fld(nhz, z.Hi[2] + o);
      if (first_row) {
        fld(nexu, z.Ei[0] + o - sy);
        fld(nezu, z.Ei[2] + o - sy);
      } else {
        fzero(nexu);
        fzero(nezu);
      }
      if (pml) load_qh(k + 1, nqh);
    } else {
      fzero(nex);
      fzero(ney);
      fzero(nez);
      fzero(nhx);
      fzero(nhy);
      fzero(nhz);
      fzero(nexu);
      fzero(nezu);
#pragma unroll
      for (int t = 0; t < 6; ++t) fzero(nqh[t]);
    }
    // (3) share E(n) of plane k
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      le[0][w][x0 + e] = ex[e];
      le[1][w][x0 + e] = ey[e];
      le[2][w][x0 + e] = ez[e];
    }
    __syncthreads();
    // (4) H(n+1/2) on plane k
    const bool store_h = own && k < ke;
    const int szs = (pml && row_ok) ? fslab(k, nz, g.pml[4], g.pml[5]) : -1;
    if (row_ok) {
      const bool pml_any = lane_pml_x || sys >= 0 || szs >= 0;
      T zbhk = 1, zchk = 0, zikhk = 1;
      if (szs >= 0) {
        zbhk = zbh[k];
        zchk = zch[k];
        zikhk = zikh[k];
      }
      const T db = (T)g.db_const;
#pragma unroll
      for (int e = 0; e < VEC; ++e) {
        const int i = x0 + e;
        const T exj = w > 0 ? le[0][w - 1][i] : exu[e];
        const T ezj = w > 0 ? le[2][w - 1][i] : ezu[e];
        const T ezi = i > 0 ? le[2][w][i - 1] : (T)0;
        const T eyi = i > 0 ? le[1][w][i - 1] : (T)0;
        T c1 = (ez[e] - ezj) * idy;  // Hx: dEz/dy - dEy/dz
        T c2 = (ey[e] - eym[e]) * idz;
        T d1 = (ex[e] - exm[e]) * idz;  // Hy: dEx/dz - dEz/dx
        T d2 = (ez[e] - ezi) * idx;
        T f1 = (ey[e] - eyi) * idx;  // Hz: dEy/dx - dEx/dy
        T f2 = (ex[e] - exj) * idy;
        if (pml_any && i < nx) {
          if (sys >= 0) {
            const int64_t o = fslab_off(g, 1, sys, i, j, k);
            const T p1 = ybh * qh[3][e] + ych * c1, p0 = ybh * qh[2][e] + ych * f2;
            if (store_h && i < nx - 1) z.psh_out[1][1][o] = p1;
            if (store_h && k < nz - 1) z.psh_out[1][0][o] = p0;
            c1 = c1 * yikh + p1;
            f2 = f2 * yikh + p0;
          }
          if (szs >= 0) {
            const int64_t o = fslab_off(g, 2, szs, i, j, k);
            const T p0 = zbhk * qh[4][e] + zchk * c2, p1 = zbhk * qh[5][e] + zchk * d1;
            if (store_h && i < nx - 1) z.psh_out[2][0][o] = p0;
            if (store_h && j < ny - 1) z.psh_out[2][1][o] = p1;
            c2 = c2 * zikhk + p0;
            d1 = d1 * zikhk + p1;
          }
          if (sxs[e] >= 0) {
            const int64_t o = fslab_off(g, 0, sxs[e], i, j, k);
            const T p0 = xbh[e] * qh[0][e] + xch[e] * d2, p1 = xbh[e] * qh[1][e] + xch[e] * f1;
            if (store_h && j < ny - 1) z.psh_out[0][0][o] = p0;
            if (store_h && k < nz - 1) z.psh_out[0][1][o] = p1;
            d2 = d2 * xikh[e] + p0;
            f1 = f1 * xikh[e] + p1;
          }
        }
        if constexpr (CK == PM_COEFF_FULL) {
          hx[e] = ((const T*)g.da[0])[off + e] * hxi[e] - ((const T*)g.db[0])[off + e] * (c1 - c2);
          hy[e] = ((const T*)g.da[1])[off + e] * hyi[e] - ((const T*)g.db[1])[off + e] * (d1 - d2);
          hz[e] = ((const T*)g.da[2])[off + e] * hzi[e] - ((const T*)g.db[2])[off + e] * (f1 - f2);
        } else {
          hx[e] = hxi[e] - db * (c1 - c2);
          hy[e] = hyi[e] - db * (d1 - d2);
          hz[e] = hzi[e] - db * (f1 - f2);
        }
        if (i >= nx - 1) hx[e] = 0;
        if (i >= nx || j >= ny - 1) hy[e] = 0;
        if (i >= nx || k >= nz - 1) hz[e] = 0;
      }
      if (store_h) {
        fst(z.Ho[0] + off, hx);
        fst(z.Ho[1] + off, hy);
        fst(z.Ho[2] + off, hz);
      }
    } else {
      fzero(hx);
      fzero(hy);
      fzero(hz);
    }
    // (5) share H(n+1/2) of plane k
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      lh[k & 1][0][w][x0 + e] = hx[e];
      lh[k & 1][1][w][x0 + e] = hy[e];
      lh[k & 1][2][w][x0 + e] = hz[e];
    }
    __syncthreads();
    // (6) E(n+1) on plane k-1
    if (do_e) {
      const uint32_t o = off - sz;
      const int p = kk & 1;
      const bool pml_any = lane_pml_x || sys >= 0 || szz >= 0;
      T zbek = 1, zcek = 0, zikek = 1;
      if (szz >= 0) {
        zbek = zbe[kk];
        zcek = zce[kk];
        zikek = zike[kk];
      }
      const bool ade = g.n_ade && j >= g.ade_j0 && j < g.ade_j1 && kk >= g.ade_k0 &&
                       kk < g.ade_k1;
      T ox[VEC], oy[VEC], oz[VEC];
#pragma unroll
      for (int e = 0; e < VEC; ++e) {
        const int i = x0 + e;
        const int ir = i + 1 < RW ? i + 1 : i;
        const T hyx = lh[p][1][w][ir], hzx = lh[p][2][w][ir];
        const T hxy = lh[p][0][w + 1][i], hzy = lh[p][2][w + 1][i];
        T a1 = (hzy - hzp[e]) * idy;  // Ex: dHz/dy - dHy/dz
        T a2 = (hy[e] - hyp[e]) * idz;
        T b1 = (hx[e] - hxp[e]) * idz;  // Ey: dHx/dz - dHz/dx
        T b2 = (hzx - hzp[e]) * idx;
        T c1 = (hyx - hyp[e]) * idx;  // Ez: dHy/dx - dHx/dy
        T c2 = (hxy - hxp[e]) * idy;
        const bool exok = i < nx && j < ny - 1 && kk < nz - 1;
        const bool eyok = i < nx - 1 && kk < nz - 1;
        const bool ezok = i < nx - 1 && j < ny - 1;
        if (pml_any) {
          if (sys >= 0) {
            const int64_t oe = fslab_off(g, 1, sys, i, j, kk);
            const T p1 = ybe * qe[3][e] + yce * a1, p0 = ybe * qe[2][e] + yce * c2;
            if (exok) ((T*)g.psi_e[1][1])[oe] = p1;
            if (ezok) ((T*)g.psi_e[1][0])[oe] = p0;
            a1 = a1 * yike + p1;
            c2 = c2 * yike + p0;
          }
          if (szz >= 0) {
            const int64_t oe = fslab_off(g, 2, szz, i, j, kk);
            const T p0 = zbek * qe[4][e] + zcek * a2, p1 = zbek * qe[5][e] + zcek * b1;
            if (exok) ((T*)g.psi_e[2][0])[oe] = p0;
            if (eyok) ((T*)g.psi_e[2][1])[oe] = p1;
            a2 = a2 * zikek + p0;
            b1 = b1 * zikek + p1;
          }
          if (sxs[e] >= 0) {
            const int64_t oe = fslab_off(g, 0, sxs[e], i, j, kk);
            const T p0 = xbe[e] * qe[0][e] + xce[e] * b2, p1 = xbe[e] * qe[1][e] + xce[e] * c1;
            if (eyok) ((T*)g.psi_e[0][0])[oe] = p0;
            if (ezok) ((T*)g.psi_e[0][1])[oe] = p1;
            b2 = b2 * xike[e] + p0;
            c1 = c1 * xike[e] + p1;
          }
        }
        const T cx = a1 - a2, cy = b1 - b2, cz = c1 - c2;
        T vx = exm[e], vy = eym[e], vz = ezm[e];
        if constexpr (CK == PM_COEFF_FULL) {
          vx = ((const T*)g.ca[0])[o + e] * vx + cbx[e] * cx;
          vy = ((const T*)g.ca[1])[o + e] * vy + cby[e] * cy;
          vz = ((const T*)g.ca[2])[o + e] * vz + cbz[e] * cz;
        } else {
          vx = vx + cbx[e] * cx;
          vy = vy + cby[e] * cy;
          vz = vz + cbz[e] * cz;
        }
        if (ade) {
          if (exok) {
            const AdeOut<T> r = fade<T>(g.ade, g.n_ade, 0, i, j, kk, cx, exm[e]);
            if (r.hit) vx = r.e;
          }
          if (eyok) {
            const AdeOut<T> r = fade<T>(g.ade, g.n_ade, 1, i, j, kk, cy, eym[e]);
            if (r.hit) vy = r.e;
          }
          if (ezok) {
            const AdeOut<T> r = fade<T>(g.ade, g.n_ade, 2, i, j, kk, cz, ezm[e]);
            if (r.hit) vz = r.e;
          }
        }
        ox[e] = exok ? vx : (T)0;
        oy[e] = eyok ? vy : (T)0;
        oz[e] = ezok ? vz : (T)0;
      }
      fst(z.Eo[0] + o, ox);
      fst(z.Eo[1] + o, oy);
      fst(z.Eo[2] + o, oz);
    }
    // (7) rotate planes, adopt the prefetched plane
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      exm[e] = ex[e];
      eym[e] = ey[e];
      ezm[e] = ez[e];
      hxp[e] = hx[e];
      hyp[e] = hy[e];
      hzp[e] = hz[e];
      ex[e] = nex[e];
      ey[e] = ney[e];
      ez[e] = nez[e];
      hxi[e] = nhx[e];
      hyi[e] = nhy[e];
      hzi[e] = nhz[e];
      exu[e] = nexu[e];
      ezu[e] = nezu[e];
#pragma unroll
      for (int t = 0; t < 6; ++t) qh[t][e] = nqh[t][e];
    }
  }
  // top plane of the grid: only Ez exists there, no z differences
  if (ke == nz && own) {
    const int kk = nz - 1;
    const uint32_t o = rowoff + (uint32_t)kk * sz;
    const int p = kk & 1;
    T cbz[VEC];
    fcb<T, CK, VEC>(g, 2, o, cbz);
    const bool ade = g.n_ade && j >= g.ade_j0 && j < g.ade_j1 && kk >= g.ade_k0 &&
                     kk < g.ade_k1;
    T oz[VEC], zero[VEC];
    fzero(zero);
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      const int i = x0 + e;
      const int ir = i + 1 < RW ? i + 1 : i;
      T c1 = (lh[p][1][w][ir] - hyp[e]) * idx;
      T c2 = (lh[p][0][w + 1][i] - hxp[e]) * idy;
      const bool ezok = i < nx - 1 && j < ny - 1;
      if (ezok && pml) {
        if (sys >= 0) c2 = fcpml_e(g, c2, 1, 0, j, sys, i, j, kk);
        if (sxs[e] >= 0) c1 = fcpml_e(g, c1, 0, 1, i, sxs[e], i, j, kk);
      }
      const T cz = c1 - c2;
      T vz;
      if constexpr (CK == PM_COEFF_FULL) vz = ((const T*)g.ca[2])[o + e] * ezm[e] + cbz[e] * cz;
      else vz = ezm[e] + cbz[e] * cz;
      if (ade && ezok) {
        const AdeOut<T> r = fade<T>(g.ade, g.n_ade, 2, i, j, kk, cz, ezm[e]);
        if (r.hit) vz = r.e;
      }
      oz[e] = ezok ? vz : (T)0;
    }
    fst(z.Eo[0] + o, zero);
    fst(z.Eo[1] + o, zero);
    fst(z.Eo[2] + o, oz);
  }
}

template <typename T, int VEC, int W, int TY>
constexpr bool fused_fits() {
  return 64 * W * (TY + 1) <= 1024 && (size_t)9 * (TY + 1) * 64 * VEC * W * sizeof(T) <= 65536;
}

template <typename T, int CK, int VEC, int W, int TY>
int launch_fused_one(const KGrid& g, const Fz<T>& z, int kchunk, int order, hipStream_t s) {
  if constexpr (fused_fits<T, VEC, W, TY>()) {
    const int ntile_y = (g.ny + TY - 1) / TY;
    const int nchunk = (g.nz + kchunk - 1) / kchunk;
    const int nb = ntile_y * nchunk;
    const int nlaunch = ((nb + 7) / 8) * 8;
    hipLaunchKernelGGL((k_yee_fused<T, CK, VEC, W, TY>), dim3(nlaunch), dim3(64 * W * (TY + 1)),
                       0, s, g, z, kchunk, ntile_y, nchunk, nb, order);
    return 0;
  } else {
    return -1;
  }
}

template <typename T, int CK, int VEC, int W>
int launch_fused_w(const KGrid& g, const Fz<T>& z, int ty, int kchunk, int order,
                   hipStream_t s) {
  if (ty == 2) return launch_fused_one<T, CK, VEC, W, 2>(g, z, kchunk, order, s);
  if (ty == 4) return launch_fused_one<T, CK, VEC, W, 4>(g, z, kchunk, order, s);
  return -1;
}

// (vec, w) pairs with 64*vec*w = row width 128, 256 or 512
template <typename T, int CK>
int launch_fused_ck(const KGrid& g, const Fz<T>& z, int vec, int w, int ty, int kchunk,
                    int order, hipStream_t s) {
  if (vec == 1 && w == 2) return launch_fused_w<T, CK, 1, 2>(g, z, ty, kchunk, order, s);
  if (vec == 2 && w == 1) return launch_fused_w<T, CK, 2, 1>(g, z, ty, kchunk, order, s);
  if (vec == 1 && w == 4) return launch_fused_w<T, CK, 1, 4>(g, z, ty, kchunk, order, s);
  if (vec == 2 && w == 2) return launch_fused_w<T, CK, 2, 2>(g, z, ty, kchunk, order, s);
  if (vec == 4 && w == 1) return launch_fused_w<T, CK, 4, 1>(g, z, ty, kchunk, order, s);
  if (vec == 2 && w == 4) return launch_fused_w<T, CK, 2, 4>(g, z, ty, kchunk, order, s);
  if (vec == 4 && w == 2) return launch_fused_w<T, CK, 4, 2>(g, z, ty, kchunk, order, s);
  return -1;
}

template <typename T>
Fz<T> make_fz(const FusedBufs& fb) {
  Fz<T> z{};
  for (int a = 0; a < 3; ++a) {
    z.Ei[a] = (const T*)fb.Ei[a];
    z.Hi[a] = (const T*)fb.Hi[a];
    z.Eo[a] = (T*)fb.Eo[a];
    z.Ho[a] = (T*)fb.Ho[a];
    for (int q = 0; q < 2; ++q) {
      z.psh_in[a][q] = (const T*)fb.psh_in[a][q];
      z.psh_out[a][q] = (T*)fb.psh_out[a][q];
    }
  }
  return z;
}

}  // namespace pm
