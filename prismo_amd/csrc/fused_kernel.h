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
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <type_traits>

#include "pm_internal.h"

namespace pm {

// Every array the fused kernel touches, as 32-bit byte offsets from one of
// three base addresses: the field set read this step (bi), the set written
// (bo), and everything else (bc: coefficients, CPML memory and profiles).
// The launcher checks each group spans less than 4 GiB.  One SGPR per array
// instead of a pointer pair keeps the loop's uniform state in registers.
struct KFz {
  const char* bi;
  const char* bo;
  const char* bc;
  uint32_t cs;      // bytes from one field component to the next (Ex Ey Ez Hx Hy Hz)
  uint32_t cb, ccs; // Cb of Ex; bytes between the three Cb arrays
  uint32_t ca[3], da[3], db[3];
  uint32_t pe[3], phi[3], pho[3];  // CPML memory per axis: q = 0 then q = 1
  uint32_t peo[3];                 // E-side memory written (two-step pass; else = pe)
  uint32_t ps[3];                  // bytes of one CPML memory array of axis a
  uint32_t prof[3];                // per axis: be ce ike bh ch ikh, n_a each
  uint32_t pn[3];                  // bytes of one profile array of axis a
};

template <typename T>
__device__ __forceinline__ const T* zp(const char* base, uint32_t off, uint32_t elem) {
  return (const T*)(base + (off + elem * (uint32_t)sizeof(T)));
}

template <typename T>
__device__ __forceinline__ T* zpw(const char* base, uint32_t off, uint32_t elem) {
  return (T*)(base + (off + elem * (uint32_t)sizeof(T)));
}

__device__ __forceinline__ int fslab(int p, int n, int tlo, int thi) {
  if (p < tlo) return p;
  if (p >= n - thi) return tlo + (p - (n - thi));
  return -1;
}

__device__ __forceinline__ uint32_t fslab_off(const KGrid& g, int a, int s, int i, int j, int k) {
  if (a == 0) {
    const uint32_t S = (uint32_t)(g.pml[0] + g.pml[1]);
    return (uint32_t)s + S * ((uint32_t)j + (uint32_t)g.ny * (uint32_t)k);
  } else if (a == 1) {
    const uint32_t S = (uint32_t)(g.pml[2] + g.pml[3]);
    return (uint32_t)i + (uint32_t)g.nx * ((uint32_t)s + S * (uint32_t)k);
  }
  return (uint32_t)i + (uint32_t)g.nx * ((uint32_t)j + (uint32_t)g.ny * (uint32_t)s);
}

// Lane l's value of v, for every lane (v_readlane; FP64 as two halves).
template <typename T>
__device__ __forceinline__ T lane_bcast(T v, int l) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
  } else {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
  }
}

// E-side CPML term, in place (E is computed on own cells only).
template <typename T>
__device__ __forceinline__ T fcpml_e(const KGrid& g, const KFz& z, T d, int a, int q, int p,
                                     int s, int i, int j, int k) {
  T* psi = zpw<T>(z.bc, z.pe[a] + (uint32_t)(q) * z.ps[a], fslab_off(g, a, s, i, j, k));
  const T ps = *zp<T>(z.bc, z.prof[a] + 0u * z.pn[a], p) * (*psi) + *zp<T>(z.bc, z.prof[a] + 1u * z.pn[a], p) * d;
  *psi = ps;
  return d * *zp<T>(z.bc, z.prof[a] + 2u * z.pn[a], p) + ps;
}

// Dispersive E update from the device box table (see yee.hip for the math).
template <typename T>
struct AdeOut {
  T e;
  bool hit;
};

template <typename T>
__device__ __forceinline__ AdeOut<T> fade(const KAde* tab, int n_ade, int comp, int i, int j,
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

// Linear index of cell (i, j, k) in box a's state arrays (box order, i fastest).
__device__ __forceinline__ int64_t ade_index(const KAde& a, int i, int j, int k) {
  return (int64_t)(i - a.i0) +
         (int64_t)(a.i1 - a.i0) * ((int64_t)(j - a.j0) + (int64_t)(a.j1 - a.j0) * (k - a.k0));
}

// Does row j of plane k cut any dispersive box? (same for every lane)  The
// union of the boxes' rows and planes (kernel arguments) rules most rows out
// before the box table (the workgroup's LDS copy) is read.
__device__ __forceinline__ bool row_has_ade(const KGrid& g, const KAde* tab, int j, int k) {
  if (j < g.ade_j0 || j >= g.ade_j1 || k < g.ade_k0 || k >= g.ade_k1) return false;
  for (int b = 0; b < g.n_ade; ++b) {
    const KAde& a = tab[b];
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

// Stores of the output field set are non-temporal: nothing reads them before
// the next step, a whole field set of traffic later, so they should not
// displace the read-only coefficients from the caches.
template <typename T, int VEC>
__device__ __forceinline__ void fst(T* p, const T (&s)[VEC]) {
  if constexpr (VEC == 1) {
    __builtin_nontemporal_store(s[0], p);
  } else {
    using VT = typename FVec<T, VEC>::type;
    VT v;
#pragma unroll
    for (int e = 0; e < VEC; ++e) v[e] = s[e];
    __builtin_nontemporal_store(v, (VT*)p);
  }
}

template <typename T, int VEC>
__device__ __forceinline__ void fzero(T (&d)[VEC]) {
#pragma unroll
  for (int e = 0; e < VEC; ++e) d[e] = (T)0;
}

// Cb as stored (FP16 or the field type), loaded early and converted where
// it is used, so the load's latency is not waited for right away.
template <typename T, int CK>
using CbRaw = typename std::conditional<CK == PM_COEFF_CB_F16, _Float16, T>::type;

template <typename T, int CK, int VEC>
__device__ __forceinline__ void fcb(const KFz& z, int c, uint32_t off, CbRaw<T, CK> (&cb)[VEC]) {
  using R = CbRaw<T, CK>;
  const R* h = (const R*)(z.bc + (z.cb + (uint32_t)c * z.ccs + off * (uint32_t)sizeof(R)));
  if constexpr (VEC == 1) {
    cb[0] = h[0];
  } else {
    using HV = R __attribute__((ext_vector_type(VEC)));
    const HV v = *(const HV*)h;
#pragma unroll
    for (int e = 0; e < VEC; ++e) cb[e] = v[e];
  }
}

template <typename T, int CK>
__device__ __forceinline__ T fcbv(const KGrid& g, CbRaw<T, CK> r) {
  if constexpr (CK == PM_COEFF_CB_F16) return (T)g.cb_scale * (T)r;
  else return r;
}

// Tiles whose rows and planes all lie outside the y and z CPML slabs form a
// box [ty0, ty1) x [kc0, kc1) of (y-tile, z-chunk); they run in a kernel
// instance built without the y/z CPML code (YZ = false), the rest in one
// with it.  Keeping the rarely used slab state out of the common instance
// keeps its uniform values in SGPRs.
struct TileBox {
  int ty0, ty1, kc0, kc1;
};

template <typename T, int VEC, int W, int TY, int XS>
struct FusedLds {
  static constexpr int RW = 64 * VEC * W;
  T le[3][TY + 1][RW];     // E(n) of plane k
  T lh[2][3][TY + 1][RW];  // H(n+1/2) of planes with parity 0 / 1
  T xr[XS > 1 ? TY + 1 : 1];  // Ex(n) of plane k one cell right of the segment
};

// One tile (TY rows x one z-chunk x one row segment of RW cells); b = the
// workgroup's index within the launch segment of its kind (slab tiles or
// interior tiles).
//
// Row segments (XS > 1: a padded row of XS * RW cells in XS workgroups).
// The H update of a segment's first cell takes E(n) of the cell to its left
// from memory (E(n) is only read this step), and the E update of its last
// cell takes H(n+1/2) of the cell to its right, which the last lane
// recomputes from E(n), H(n-1/2) and the CPML memory of that cell with the
// arithmetic the neighbouring segment uses for it (not stored).  Segment
// boundaries lie outside the x CPML layers (the host checks).
template <typename T, int CK, int VEC, int W, int TY, int XS, bool YZ, bool ADE>
__device__ __forceinline__ void fused_tile(const KGrid& g, const KFz& z, int kchunk,
                                           int ntile_y, int nchunk, int nblocks, int order,
                                           const TileBox& bx, int b,
                                           FusedLds<T, VEC, W, TY, XS>& lds, const KAde* tab,
                                           const T (*tcf)[12]) {
  constexpr int RW = 64 * VEC * W;
  using ABits = std::conditional_t<(12 * VEC > 32), uint64_t, uint32_t>;
  static_assert(kMaxAde < 16, "box numbers are kept in 4 bits");
  auto& le = lds.le;
  auto& lh = lds.lh;
  const int per = (nblocks + 7) / 8;
  // orders 0/1: XCD b % 8 runs a contiguous range of tiles (halo reuse in its
  // L2); order 2: hardware round-robin (slab tiles spread over all XCDs);
  // order 3: the tile list in groups of 64, XCD x taking tiles 8x .. 8x+7 of
  // each group, so the XCDs advance through the grid together while each
  // runs runs of 8 neighbouring tiles (a remainder under 64 goes round-robin)
  int L;
  if (order == 2) {
    L = b;
  } else if (order == 3) {
    L = b < (nblocks & ~63) ? (b & ~63) + (b & 7) * 8 + ((b >> 3) & 7) : b;
  } else {
    L = (b % 8) * per + b / 8;
  }
  if (L >= nblocks) return;
  // consecutive tile numbers are the segments of one row tile
  const int xs = XS > 1 ? L % XS : 0;
  if constexpr (XS > 1) L /= XS;
  int ty, kc;
  if constexpr (!YZ) {
    const int nty = bx.ty1 - bx.ty0, nkc = bx.kc1 - bx.kc0;
    if (order != 1) {
      ty = bx.ty0 + L % nty;
      kc = bx.kc0 + L / nty;
    } else {
      kc = bx.kc0 + L % nkc;
      ty = bx.ty0 + L / nkc;
    }
  } else {
    // every tile outside the box: chunks below it, chunks above it, then
    // the y-tiles beside it in the chunks it spans
    const int n_lo = bx.kc0 * ntile_y;
    const int n_hi = (nchunk - bx.kc1) * ntile_y;
    if (L < n_lo) {
      kc = L / ntile_y;
      ty = L % ntile_y;
    } else if (L < n_lo + n_hi) {
      kc = bx.kc1 + (L - n_lo) / ntile_y;
      ty = (L - n_lo) % ntile_y;
    } else {
      const int nb = ntile_y - (bx.ty1 - bx.ty0);
      const int r = (L - n_lo - n_hi) % nb;
      kc = bx.kc0 + (L - n_lo - n_hi) / nb;
      ty = r < bx.ty0 ? r : bx.ty1 + (r - bx.ty0);
    }
  }

  const int tid = threadIdx.x;
  // the row index is the same for a whole wave: make it scalar so row-wise
  // quantities (y profiles, slab indices) live in SGPRs and load as scalars
  const int w = __builtin_amdgcn_readfirstlane(tid / (64 * W));  // TY = halo row
  const int x0 = (tid % (64 * W)) * VEC;  // within the segment (LDS)
  const int gx0 = xs * RW + x0;             // within the row (memory)
  // this segment has a segment to its left / right
  const bool seg_l = XS > 1 && xs > 0;
  const bool seg_r = XS > 1 && xs + 1 < XS;
  const bool lane_r = x0 == RW - VEC;
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
  const bool first_row = (w == 0) && row_ok && j > 0;
  const bool pml = g.has_pml;

  // ---- CPML profiles: x per lane (registers), y per row (scalars) -------
  const int sys = (YZ && pml && row_ok) ? fslab(j, ny, g.pml[2], g.pml[3]) : -1;
  int sxs[VEC];
  bool lane_pml_x = false;
  T xbh[VEC], xch[VEC], xikh[VEC], xbe[VEC], xce[VEC], xike[VEC];
#pragma unroll
  for (int e = 0; e < VEC; ++e) {
    const int i = gx0 + e;
    sxs[e] = (pml && i < nx) ? fslab(i, nx, g.pml[0], g.pml[1]) : -1;
    lane_pml_x |= sxs[e] >= 0;
    xbh[e] = xbe[e] = (T)1;
    xch[e] = xce[e] = (T)0;
    xikh[e] = xike[e] = (T)1;
    if (sxs[e] >= 0) {
      xbh[e] = *zp<T>(z.bc, z.prof[0] + 3u * z.pn[0], i);
      xch[e] = *zp<T>(z.bc, z.prof[0] + 4u * z.pn[0], i);
      xikh[e] = *zp<T>(z.bc, z.prof[0] + 5u * z.pn[0], i);
      xbe[e] = *zp<T>(z.bc, z.prof[0] + 0u * z.pn[0], i);
      xce[e] = *zp<T>(z.bc, z.prof[0] + 1u * z.pn[0], i);
      xike[e] = *zp<T>(z.bc, z.prof[0] + 2u * z.pn[0], i);
    }
  }
  T ybh = 1, ych = 0, yikh = 1, ybe = 1, yce = 0, yike = 1;
  if (sys >= 0) {
    ybh = *zp<T>(z.bc, z.prof[1] + 3u * z.pn[1], j);
    ych = *zp<T>(z.bc, z.prof[1] + 4u * z.pn[1], j);
    yikh = *zp<T>(z.bc, z.prof[1] + 5u * z.pn[1], j);
    ybe = *zp<T>(z.bc, z.prof[1] + 0u * z.pn[1], j);
    yce = *zp<T>(z.bc, z.prof[1] + 1u * z.pn[1], j);
    yike = *zp<T>(z.bc, z.prof[1] + 2u * z.pn[1], j);
  }

  // Every global load in the z-march is issued on every path, from a
  // clamped in-range position where its value goes unused (rows past ny,
  // the plane after the chunk, lanes outside a CPML slab).  The compiler
  // then counts the same number of loads on every path, so each memory wait
  // covers only the loads it needs and never the plane still in flight.
  // Planes alternate between two register sets (the loop is unrolled by
  // two), so no in-flight register is ever copied.
  const int jr = min(j, ny - 1);
  const uint32_t rowoff = (uint32_t)jr * sy + (uint32_t)gx0;
  const uint32_t upoff = (uint32_t)(first_row ? j - 1 : jr) * sy + (uint32_t)gx0;
  // the halo row takes no E update: its Cb loads read the row above, which
  // the tile's last own row fetches at the same time (a cache hit)
  const uint32_t cboff = (uint32_t)(w < TY ? jr : min(j - 1, ny - 1)) * sy + (uint32_t)gx0;
  int sxc[VEC];
#pragma unroll
  for (int e = 0; e < VEC; ++e) sxc[e] = max(sxs[e], 0);
  const int syc = max(sys, 0);

  // CPML memory of one plane, [x0 x1 y0 y1 z0 z1] per cell; base = the
  // first array of each axis' pair (an axis without CPML points at a large
  // array the host picks, so the clamped load stays in bounds)
  auto load_psi = [&](const uint32_t(&base)[3], int kq, T(&q)[6][VEC])
                      __attribute__((always_inline)) {
    const int skc = YZ ? max(fslab(kq, nz, g.pml[4], g.pml[5]), 0) : 0;
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      const int ic = min(gx0 + e, nx - 1);
      const uint32_t ox = fslab_off(g, 0, sxc[e], ic, jr, kq);
      q[0][e] = *zp<T>(z.bc, base[0], ox);
      q[1][e] = *zp<T>(z.bc, base[0] + z.ps[0], ox);
      if constexpr (YZ) {
        const uint32_t oy = fslab_off(g, 1, syc, ic, jr, kq);
        q[2][e] = *zp<T>(z.bc, base[1], oy);
        q[3][e] = *zp<T>(z.bc, base[1] + z.ps[1], oy);
        const uint32_t oz = fslab_off(g, 2, skc, ic, jr, kq);
        q[4][e] = *zp<T>(z.bc, base[2], oz);
        q[5][e] = *zp<T>(z.bc, base[2] + z.ps[2], oz);
      } else {
#pragma unroll
        for (int t = 2; t < 6; ++t) q[t][e] = (T)0;
      }
    }
  };

  // Row segments: the values a segment needs from beyond its ends, one per
  // lane of every wave, loaded with the plane (lanes 0-1: Ey, Ez left of
  // the segment; 2-6: Ex, Ey, Ez, Hy, Hz right of it; 7: Ex right of it on
  // the row above the tile's first row; 8-9, slab tiles: the y and z H-side
  // CPML memory right of it; other lanes repeat lane 2).  Scalar loads
  // would cost no vector registers, but their counter is the one LDS
  // accesses use and they return out of order, so the first LDS wait of
  // every plane would wait for them.  A segment at the row's end reads
  // column 0 or nx-1 (unused).  FP64 instances keep the scalar loads: at
  // 168 VGPRs the lane-spread load's two registers per plane set cost them
  // spills.
  constexpr bool HL = XS > 1 && sizeof(T) == 4;
  const int hlane = tid & 63;
  const char* hbase = z.bi;
  uint32_t hoff0 = 0, hstep = 0;
  if constexpr (HL) {
    const uint32_t esz = (uint32_t)sizeof(T);
    const uint32_t rcol0 = (uint32_t)(seg_r ? min((xs + 1) * RW, nx - 1) : 0);
    const uint32_t lcol0 = (uint32_t)(seg_l ? xs * RW - 1 : 0);
    const uint32_t rowr = (uint32_t)jr * sy + rcol0;
    hstep = sz * esz;
    if (hlane < 2) {
      hoff0 = (uint32_t)(hlane + 1) * z.cs + ((uint32_t)jr * sy + lcol0) * esz;
    } else if (hlane < 7) {
      const uint32_t c = hlane == 2 ? 0u : hlane == 3 ? 1u : hlane == 4 ? 2u : hlane == 5 ? 4u : 5u;
      hoff0 = c * z.cs + rowr * esz;
    } else if (hlane == 7) {
      hoff0 = ((uint32_t)(first_row ? j - 1 : jr) * sy + rcol0) * esz;
    } else if (YZ && hlane == 8) {
      hbase = z.bc;
      const uint32_t Sy = (uint32_t)(g.pml[2] + g.pml[3]);
      hoff0 = z.phi[1] + fslab_off(g, 1, max(sys, 0), (int)rcol0, jr, 0) * esz;
      hstep = (uint32_t)nx * Sy * esz;
    } else if (YZ && hlane == 9) {
      hbase = z.bc;
      hoff0 = z.phi[2] + z.ps[2] + (rcol0 + (uint32_t)nx * (uint32_t)jr) * esz;
      hstep = 0;
    } else {
      hoff0 = rowr * esz;
    }
  }

  struct PlaneIn {
    T ex[VEC], ey[VEC], ez[VEC];  // E(n) of the plane
    T hx[VEC], hy[VEC], hz[VEC];  // H(n-1/2)
    T exu[VEC], ezu[VEC];         // row j-1 (first row of the tile only)
    T q[6][VEC];                  // H-side CPML memory
    T hl;                         // row segments: this lane's value from beyond the ends
  };
  auto load_plane = [&](int kq, PlaneIn& P) __attribute__((always_inline)) {
    const uint32_t o = rowoff + (uint32_t)kq * sz;
    fld(P.ex, zp<T>(z.bi, 0u, o));
    fld(P.ey, zp<T>(z.bi, z.cs, o));
    fld(P.ez, zp<T>(z.bi, 2u * z.cs, o));
    fld(P.hx, zp<T>(z.bi, 3u * z.cs, o));
    // the halo row's Hy feeds nothing (E of the last own row takes Hx and Hz
    // from it): read the row above instead, a cache hit
    fld(P.hy, zp<T>(z.bi, 4u * z.cs, w < TY ? o : cboff + (uint32_t)kq * sz));
    fld(P.hz, zp<T>(z.bi, 5u * z.cs, o));
    const uint32_t u = upoff + (uint32_t)kq * sz;
    fld(P.exu, zp<T>(z.bi, 0u, u));
    fld(P.ezu, zp<T>(z.bi, 2u * z.cs, u));
    load_psi(z.phi, kq, P.q);
    if constexpr (HL) {
      uint32_t ho = hoff0 + (uint32_t)kq * hstep;
      if constexpr (YZ) {
        // z slab memory: slab plane index of kq (0 outside the z layers)
        const int skh = max(fslab(kq, nz, g.pml[4], g.pml[5]), 0);
        if (hlane == 9) ho += (uint32_t)nx * (uint32_t)ny * (uint32_t)skh * (uint32_t)sizeof(T);
      }
      P.hl = *(gptr<const T>)(hbase + ho);
    }
  };

  T exm[VEC], eym[VEC], ezm[VEC];  // E(n) of plane k-1
  T hxp[VEC], hyp[VEC], hzp[VEC];  // H(n+1/2) of plane k-1
  fzero(exm);
  fzero(eym);
  fzero(ezm);
  fzero(hxp);
  fzero(hyp);
  fzero(hzp);
  if (row_ok && kb > 0) {
    const uint32_t o = rowoff + (uint32_t)(kb - 1) * sz;
    fld(exm, zp<T>(z.bi, 0u, o));
    fld(eym, zp<T>(z.bi, z.cs, o));
  }
  // row segments, last lane: Ex(n) of the cell right of the segment on the
  // previous plane, and that cell's recomputed H(n+1/2) on the previous
  // plane (which the E update of that plane takes)
  const int gxr = (xs + 1) * RW;  // the cell right of the segment
  // the halo cells are the same for every lane of a row: their loads are
  // scalar loads (the input field set and the H-side CPML memory are not
  // written by this kernel), so they hold no vector registers
  const uint32_t rcol = (uint32_t)(seg_r ? min(gxr, nx - 1) : 0);
  const uint32_t lcol = (uint32_t)(seg_l ? xs * RW - 1 : 0);
  using CT = const T __attribute__((address_space(4)))*;
  auto sld = [&](const char* base, uint32_t off, uint32_t elem) __attribute__((always_inline)) {
    const uint32_t o =
        (uint32_t)__builtin_amdgcn_readfirstlane((int)(off + elem * (uint32_t)sizeof(T)));
    return *(CT)(base + o);
  };
  T sr_exm = 0, hyh_p = 0, hzh_p = 0;
  if constexpr (XS > 1) {
    if (row_ok && kb > 0) sr_exm = sld(z.bi, 0u, (uint32_t)jr * sy + rcol + (uint32_t)(kb - 1) * sz);
  }

  // dispersive boxes this wave's row cuts, per component (uniform), and per
  // lane those whose x range holds cell e: the z-march changes only the
  // plane, so each plane then needs the boxes' plane ranges alone
  uint32_t rc0 = 0, rc1 = 0, rc2 = 0;
  uint32_t xm[VEC];
#pragma unroll
  for (int e = 0; e < VEC; ++e) xm[e] = 0;
  if constexpr (ADE) {
    if (row_ok && j >= g.ade_j0 && j < g.ade_j1) {
      for (int t = 0; t < g.n_ade; ++t) {
        const KAde& a = tab[t];
        if (j < a.j0 || j >= a.j1) continue;
        const uint32_t bit = 1u << t;
        if (a.comp == 0)
          rc0 |= bit;
        else if (a.comp == 1)
          rc1 |= bit;
        else
          rc2 |= bit;
#pragma unroll
        for (int e = 0; e < VEC; ++e)
          if (gx0 + e >= a.i0 && gx0 + e < a.i1) xm[e] |= bit;
      }
    }
  }
  // the boxes that hold a cell of this wave (over all its lanes) and the
  // union of their plane ranges: waves with no cell in a box, and planes
  // outside that union, skip the per-plane box code (a filter only: where
  // it passes, each lane still finds its box per plane).  FP64 instances
  // filter by the row's boxes and the union of all boxes' planes instead
  // (at 168 VGPRs the per-wave filter costs them spills).
  int wk0 = g.ade_k0, wk1 = (rc0 | rc1 | rc2) != 0 ? g.ade_k1 : 0;
  if constexpr (ADE && sizeof(T) == 4) {
    uint32_t wb = 0;
#pragma unroll
    for (int e = 0; e < VEC; ++e) wb |= xm[e];
    for (int m = 32; m >= 1; m >>= 1) wb |= (uint32_t)__shfl_xor((int)wb, m);
    wb = (uint32_t)__builtin_amdgcn_readfirstlane((int)wb);
    wk0 = nz;
    for (int t = 0; t < g.n_ade; ++t)
      if ((wb >> t) & 1u) {
        wk0 = min(wk0, tab[t].k0);
        wk1 = max(wk1, tab[t].k1);
      }
  }

  auto step = [&](const int k, const PlaneIn& cur, PlaneIn& nxt) __attribute__((always_inline)) {
    const uint32_t off = rowoff + (uint32_t)k * sz;
    const bool do_e = k > kb && own;
    const int kk = k - 1;
    const int kq = k > kb ? kk : k;  // where the E-side loads read
    // (0) row segments: E(n) left of the segment's first cell, and E(n),
    //     H(n-1/2) and the CPML memory of the cell right of its last one
    //     (scalar loads, one per row; a segment at the row's end reads
    //     column 0 or nx-1, unused)
    T sl_ey = 0, sl_ez = 0;
    T sr_ex = 0, sr_ey = 0, sr_ez = 0, sr_hy = 0, sr_hz = 0, sr_exu = 0, sr_q2 = 0, sr_q5 = 0;
    if constexpr (HL) {
      sl_ey = lane_bcast(cur.hl, 0);
      sl_ez = lane_bcast(cur.hl, 1);
      sr_ex = lane_bcast(cur.hl, 2);
      sr_ey = lane_bcast(cur.hl, 3);
      sr_ez = lane_bcast(cur.hl, 4);
      sr_hy = lane_bcast(cur.hl, 5);
      sr_hz = lane_bcast(cur.hl, 6);
      sr_exu = lane_bcast(cur.hl, 7);
      if constexpr (YZ) {
        sr_q2 = lane_bcast(cur.hl, 8);
        sr_q5 = lane_bcast(cur.hl, 9);
      }
    } else if constexpr (XS > 1) {
      const uint32_t lo = (uint32_t)jr * sy + lcol + (uint32_t)k * sz;
      sl_ey = sld(z.bi, z.cs, lo);
      sl_ez = sld(z.bi, 2u * z.cs, lo);
      const uint32_t ro = (uint32_t)jr * sy + rcol + (uint32_t)k * sz;
      sr_ex = sld(z.bi, 0u, ro);
      sr_ey = sld(z.bi, z.cs, ro);
      sr_ez = sld(z.bi, 2u * z.cs, ro);
      sr_hy = sld(z.bi, 4u * z.cs, ro);
      sr_hz = sld(z.bi, 5u * z.cs, ro);
      sr_exu = sld(z.bi, 0u, (uint32_t)(first_row ? j - 1 : jr) * sy + rcol + (uint32_t)k * sz);
      if constexpr (YZ) {
        const int skh = max(fslab(k, nz, g.pml[4], g.pml[5]), 0);
        sr_q2 = sld(z.bc, z.phi[1], fslab_off(g, 1, syc, (int)rcol, jr, k));
        sr_q5 = sld(z.bc, z.phi[2] + z.ps[2], fslab_off(g, 2, skh, (int)rcol, jr, k));
      }
    }
    // (1) what this iteration's E update needs, issued ahead of the next
    //     plane's prefetch: vector memory counters retire in order, so the
    //     wait for these does not wait for the prefetch
    CbRaw<T, CK> cbx[VEC], cby[VEC], cbz[VEC];
    const uint32_t oq = cboff + (uint32_t)kq * sz;
    fcb<T, CK, VEC>(z, 0, oq, cbx);
    fcb<T, CK, VEC>(z, 1, oq, cby);
    fcb<T, CK, VEC>(z, 2, oq, cbz);
    T qe[6][VEC];
    load_psi(z.pe, kq, qe);
    // z profiles of plane k (H side) and of the E plane (slab tiles only)
    T zh[3] = {1, 0, 1}, ze[3] = {1, 0, 1};
    if constexpr (YZ) {
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        zh[t] = *zp<T>(z.bc, z.prof[2] + (uint32_t)(3 + t) * z.pn[2], k);
        ze[t] = *zp<T>(z.bc, z.prof[2] + (uint32_t)t * z.pn[2], kq);
      }
    }
    // (1b) dispersive rows: the box of each lane's cell per component (the
    //      first box holding the cell takes it: the lowest set bit of the
    //      row's, the plane's and the lane's box masks), then its state (P,
    //      and Q for Lorentz), issued with the E-side loads so the E update
    //      does not wait for them.  No loop here or in (6) issues a memory
    //      access: with a loop-dependent count of accesses in flight the
    //      compiler would wait for all of them, the plane prefetch included.
    const bool ade_row = ADE && do_e && kk >= wk0 && kk < wk1;
    // box number + 1 of (component c, cell e) in 4 bits at 4 (c VEC + e);
    // 0: no box
    ABits abits = 0;
    T ap[3][VEC], aq[3][VEC];
    if constexpr (ADE) {
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int e = 0; e < VEC; ++e) ap[c][e] = aq[c][e] = (T)0;
      if (ade_row) {
        uint32_t km = 0;  // boxes whose plane range holds kk
        for (int t = 0; t < g.n_ade; ++t) {
          const KAde& a = tab[t];
          if (kk >= a.k0 && kk < a.k1) km |= 1u << t;
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const uint32_t cm = (c == 0 ? rc0 : c == 1 ? rc1 : rc2) & km;
#pragma unroll
          for (int e = 0; e < VEC; ++e) {
            const uint32_t m = cm & xm[e];
            if (m) abits |= (ABits)(__builtin_ctz(m) + 1) << (4 * (c * VEC + e));
          }
        }
        // every lane loads: one without a box reads the first word of the
        // device copy of the table (valid memory, value unused)
        const gptr<const T> dummy = (gptr<const T>)g.ade;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int e = 0; e < VEC; ++e) {
            const int t = (int)((abits >> (4 * (c * VEC + e))) & 15) - 1;
            const KAde& a = tab[t < 0 ? 0 : t];
            const uint32_t l = t >= 0 ? (uint32_t)ade_index(a, gx0 + e, j, kk) : 0u;
            const gptr<const T> pp = t >= 0 ? (gptr<const T>)a.p + l : dummy;
            const gptr<const T> qq = (t >= 0 && a.kind == 1) ? (gptr<const T>)a.q + l : dummy;
            ap[c][e] = *pp;
            aq[c][e] = *qq;
          }
      }
    }
    // (2) prefetch plane k+1 (plane kh again after the last plane: unused)
    load_plane(min(k + 1, kh), nxt);
    // (3) share E(n) of plane k
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      le[0][w][x0 + e] = cur.ex[e];
      le[1][w][x0 + e] = cur.ey[e];
      le[2][w][x0 + e] = cur.ez[e];
    }
    if constexpr (XS > 1) {
      if (lane_r) lds.xr[w] = sr_ex;
    }
    __syncthreads();
    // (4) H(n+1/2) on plane k
    T hx[VEC], hy[VEC], hz[VEC];
    const bool store_h = own && k < ke;
    const int szs = (YZ && pml && row_ok) ? fslab(k, nz, g.pml[4], g.pml[5]) : -1;
    if (row_ok) {
      const bool pml_any = lane_pml_x || sys >= 0 || szs >= 0;
      const T zbhk = zh[0], zchk = zh[1], zikhk = zh[2];
      const T db = (T)g.db_const;
#pragma unroll
      for (int e = 0; e < VEC; ++e) {
        const int li = x0 + e, i = gx0 + e;
        const T ex = cur.ex[e], ey = cur.ey[e], ez = cur.ez[e];
        const T exj = w > 0 ? le[0][w - 1][li] : (first_row ? cur.exu[e] : (T)0);
        const T ezj = w > 0 ? le[2][w - 1][li] : (first_row ? cur.ezu[e] : (T)0);
        const T ezi = li > 0 ? le[2][w][li - 1] : (seg_l ? sl_ez : (T)0);
        const T eyi = li > 0 ? le[1][w][li - 1] : (seg_l ? sl_ey : (T)0);
        T c1 = (ez - ezj) * idy;  // Hx: dEz/dy - dEy/dz
        T c2 = (ey - eym[e]) * idz;
        T d1 = (ex - exm[e]) * idz;  // Hy: dEx/dz - dEz/dx
        T d2 = (ez - ezi) * idx;
        T f1 = (ey - eyi) * idx;  // Hz: dEy/dx - dEx/dy
        T f2 = (ex - exj) * idy;
        if (pml_any && i < nx) {
          if (sys >= 0) {
            const uint32_t o = fslab_off(g, 1, sys, i, j, k);
            const T p1 = ybh * cur.q[3][e] + ych * c1, p0 = ybh * cur.q[2][e] + ych * f2;
            if (store_h && i < nx - 1) *zpw<T>(z.bc, z.pho[1] + z.ps[1], o) = p1;
            if (store_h && k < nz - 1) *zpw<T>(z.bc, z.pho[1], o) = p0;
            c1 = c1 * yikh + p1;
            f2 = f2 * yikh + p0;
          }
          if (szs >= 0) {
            const uint32_t o = fslab_off(g, 2, szs, i, j, k);
            const T p0 = zbhk * cur.q[4][e] + zchk * c2, p1 = zbhk * cur.q[5][e] + zchk * d1;
            if (store_h && i < nx - 1) *zpw<T>(z.bc, z.pho[2], o) = p0;
            if (store_h && j < ny - 1) *zpw<T>(z.bc, z.pho[2] + z.ps[2], o) = p1;
            c2 = c2 * zikhk + p0;
            d1 = d1 * zikhk + p1;
          }
          if (sxs[e] >= 0) {
            const uint32_t o = fslab_off(g, 0, sxs[e], i, j, k);
            const T p0 = xbh[e] * cur.q[0][e] + xch[e] * d2;
            const T p1 = xbh[e] * cur.q[1][e] + xch[e] * f1;
            if (store_h && j < ny - 1) *zpw<T>(z.bc, z.pho[0], o) = p0;
            if (store_h && k < nz - 1) *zpw<T>(z.bc, z.pho[0] + z.ps[0], o) = p1;
            d2 = d2 * xikh[e] + p0;
            f1 = f1 * xikh[e] + p1;
          }
        }
        if constexpr (CK == PM_COEFF_FULL) {
          hx[e] = *zp<T>(z.bc, z.da[0], off + e) * cur.hx[e] - *zp<T>(z.bc, z.db[0], off + e) * (c1 - c2);
          hy[e] = *zp<T>(z.bc, z.da[1], off + e) * cur.hy[e] - *zp<T>(z.bc, z.db[1], off + e) * (d1 - d2);
          hz[e] = *zp<T>(z.bc, z.da[2], off + e) * cur.hz[e] - *zp<T>(z.bc, z.db[2], off + e) * (f1 - f2);
        } else {
          hx[e] = cur.hx[e] - db * (c1 - c2);
          hy[e] = cur.hy[e] - db * (d1 - d2);
          hz[e] = cur.hz[e] - db * (f1 - f2);
        }
        if (i >= nx - 1) hx[e] = 0;
        if (i >= nx || j >= ny - 1) hy[e] = 0;
        if (i >= nx || k >= nz - 1) hz[e] = 0;
      }
      if (store_h) {
        fst(zpw<T>(z.bo, 3u * z.cs, off), hx);
        fst(zpw<T>(z.bo, 4u * z.cs, off), hy);
        fst(zpw<T>(z.bo, 5u * z.cs, off), hz);
      }
    } else {
      fzero(hx);
      fzero(hy);
      fzero(hz);
    }
    // (4b) row segments: H(n+1/2) of the cell right of the segment, with the
    //      arithmetic of the H update above (that cell lies outside the x
    //      CPML layers), kept in registers for the E update of the last cell
    T hyh = 0, hzh = 0;
    if constexpr (XS > 1) {
      if (seg_r && lane_r && row_ok) {
        const T db = (T)g.db_const;
        const T exj = w > 0 ? lds.xr[w - 1] : (first_row ? sr_exu : (T)0);
        T d1 = (sr_ex - sr_exm) * idz;
        T d2 = (sr_ez - cur.ez[VEC - 1]) * idx;
        T f1 = (sr_ey - cur.ey[VEC - 1]) * idx;
        T f2 = (sr_ex - exj) * idy;
        if (gxr < nx) {
          if (sys >= 0) {
            const T p0 = ybh * sr_q2 + ych * f2;
            f2 = f2 * yikh + p0;
          }
          if (szs >= 0) {
            const T p1 = zh[0] * sr_q5 + zh[1] * d1;
            d1 = d1 * zh[2] + p1;
          }
        }
        hyh = sr_hy - db * (d1 - d2);
        hzh = sr_hz - db * (f1 - f2);
        if (gxr >= nx || j >= ny - 1) hyh = 0;
        if (gxr >= nx || k >= nz - 1) hzh = 0;
      }
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
      const int szz = (YZ && pml) ? fslab(kk, nz, g.pml[4], g.pml[5]) : -1;
      const bool pml_any = lane_pml_x || sys >= 0 || szz >= 0;
      const T zbek = ze[0], zcek = ze[1], zikek = ze[2];
      T ox[VEC], oy[VEC], oz[VEC];
#pragma unroll
      for (int e = 0; e < VEC; ++e) {
        const int li = x0 + e, i = gx0 + e;
        const int ir = li + 1 < RW ? li + 1 : li;
        T hyx = lh[p][1][w][ir], hzx = lh[p][2][w][ir];
        if (seg_r && li == RW - 1) {
          hyx = hyh_p;
          hzx = hzh_p;
        }
        const T hxy = lh[p][0][w + 1][li], hzy = lh[p][2][w + 1][li];
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
            const uint32_t oe = fslab_off(g, 1, sys, i, j, kk);
            const T p1 = ybe * qe[3][e] + yce * a1, p0 = ybe * qe[2][e] + yce * c2;
            if (exok) *zpw<T>(z.bc, z.pe[1] + z.ps[1], oe) = p1;
            if (ezok) *zpw<T>(z.bc, z.pe[1], oe) = p0;
            a1 = a1 * yike + p1;
            c2 = c2 * yike + p0;
          }
          if (szz >= 0) {
            const uint32_t oe = fslab_off(g, 2, szz, i, j, kk);
            const T p0 = zbek * qe[4][e] + zcek * a2, p1 = zbek * qe[5][e] + zcek * b1;
            if (exok) *zpw<T>(z.bc, z.pe[2], oe) = p0;
            if (eyok) *zpw<T>(z.bc, z.pe[2] + z.ps[2], oe) = p1;
            a2 = a2 * zikek + p0;
            b1 = b1 * zikek + p1;
          }
          if (sxs[e] >= 0) {
            const uint32_t oe = fslab_off(g, 0, sxs[e], i, j, kk);
            const T p0 = xbe[e] * qe[0][e] + xce[e] * b2, p1 = xbe[e] * qe[1][e] + xce[e] * c1;
            if (eyok) *zpw<T>(z.bc, z.pe[0], oe) = p0;
            if (ezok) *zpw<T>(z.bc, z.pe[0] + z.ps[0], oe) = p1;
            b2 = b2 * xike[e] + p0;
            c1 = c1 * xike[e] + p1;
          }
        }
        const T cx = a1 - a2, cy = b1 - b2, cz = c1 - c2;
        T vx = exm[e], vy = eym[e], vz = ezm[e];
        if constexpr (CK == PM_COEFF_FULL) {
          vx = *zp<T>(z.bc, z.ca[0], o + e) * vx + fcbv<T, CK>(g, cbx[e]) * cx;
          vy = *zp<T>(z.bc, z.ca[1], o + e) * vy + fcbv<T, CK>(g, cby[e]) * cy;
          vz = *zp<T>(z.bc, z.ca[2], o + e) * vz + fcbv<T, CK>(g, cbz[e]) * cz;
        } else {
          vx = vx + fcbv<T, CK>(g, cbx[e]) * cx;
          vy = vy + fcbv<T, CK>(g, cby[e]) * cy;
          vz = vz + fcbv<T, CK>(g, cbz[e]) * cz;
        }
        if constexpr (ADE) {
          if (ade_row) {
            // each lane reads the coefficients of the box it found above
            const int tx = (int)((abits >> (4 * e)) & 15) - 1;
            const int ty_ = (int)((abits >> (4 * (VEC + e))) & 15) - 1;
            const int tz = (int)((abits >> (4 * (2 * VEC + e))) & 15) - 1;
            if (exok && tx >= 0) {
              const KAde& a = tab[tx];
              vx = ade_cell_coef<T>(a, tcf[tx], ade_index(a, i, j, kk), ap[0][e], aq[0][e], cx,
                                    exm[e]);
            }
            if (eyok && ty_ >= 0) {
              const KAde& a = tab[ty_];
              vy = ade_cell_coef<T>(a, tcf[ty_], ade_index(a, i, j, kk), ap[1][e], aq[1][e], cy,
                                    eym[e]);
            }
            if (ezok && tz >= 0) {
              const KAde& a = tab[tz];
              vz = ade_cell_coef<T>(a, tcf[tz], ade_index(a, i, j, kk), ap[2][e], aq[2][e], cz,
                                    ezm[e]);
            }
          }
        }
        ox[e] = exok ? vx : (T)0;
        oy[e] = eyok ? vy : (T)0;
        oz[e] = ezok ? vz : (T)0;
      }
      fst(zpw<T>(z.bo, 0u, o), ox);
      fst(zpw<T>(z.bo, z.cs, o), oy);
      fst(zpw<T>(z.bo, 2u * z.cs, o), oz);
    }
    // (7) plane k becomes plane k-1
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      exm[e] = cur.ex[e];
      eym[e] = cur.ey[e];
      ezm[e] = cur.ez[e];
      hxp[e] = hx[e];
      hyp[e] = hy[e];
      hzp[e] = hz[e];
    }
    if constexpr (XS > 1) {
      sr_exm = sr_ex;
      hyh_p = hyh;
      hzh_p = hzh;
    }
  };

  PlaneIn pa, pb;
  load_plane(kb, pa);
  for (int k = kb; k <= kh; k += 2) {
    step(k, pa, pb);
    if (k + 1 > kh) break;
    step(k + 1, pb, pa);
  }
  // top plane of the grid: only Ez exists there, no z differences
  if (ke == nz && own) {
    const int kk = nz - 1;
    const uint32_t o = rowoff + (uint32_t)kk * sz;
    const int p = kk & 1;
    CbRaw<T, CK> cbz[VEC];
    fcb<T, CK, VEC>(z, 2, o, cbz);
    const bool ade = ADE && row_has_ade(g, tab, j, kk);
    T oz[VEC], zero[VEC];
    fzero(zero);
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      const int li = x0 + e, i = gx0 + e;
      const int ir = li + 1 < RW ? li + 1 : li;
      const T hyx = (seg_r && li == RW - 1) ? hyh_p : lh[p][1][w][ir];
      T c1 = (hyx - hyp[e]) * idx;
      T c2 = (lh[p][0][w + 1][li] - hxp[e]) * idy;
      const bool ezok = i < nx - 1 && j < ny - 1;
      if (ezok && pml) {
        if (sys >= 0) c2 = fcpml_e<T>(g, z, c2, 1, 0, j, sys, i, j, kk);
        if (sxs[e] >= 0) c1 = fcpml_e<T>(g, z, c1, 0, 1, i, sxs[e], i, j, kk);
      }
      const T cz = c1 - c2;
      T vz;
      if constexpr (CK == PM_COEFF_FULL)
        vz = *zp<T>(z.bc, z.ca[2], o + e) * ezm[e] + fcbv<T, CK>(g, cbz[e]) * cz;
      else
        vz = ezm[e] + fcbv<T, CK>(g, cbz[e]) * cz;
      if (ade && ezok) {
        const AdeOut<T> r = fade<T>(tab, g.n_ade, 2, i, j, kk, cz, ezm[e]);
        if (r.hit) vz = r.e;
      }
      oz[e] = ezok ? vz : (T)0;
    }
    fst(zpw<T>(z.bo, (uint32_t)(0) * z.cs, o), zero);
    fst(zpw<T>(z.bo, (uint32_t)(1) * z.cs, o), zero);
    fst(zpw<T>(z.bo, (uint32_t)(2) * z.cs, o), oz);
  }
}

// One launch for all tiles: workgroups [0, seg) take the CPML slab tiles
// (the slower ones, dispatched first), the rest the interior box, so the
// interior tiles fill the CUs the slab tiles leave idle at the end.  seg is
// a multiple of 8, so a workgroup's XCD (blockIdx % 8) is the same in its
// segment's numbering.
template <typename T, int CK, int VEC, int W, int TY, int XS, bool ADE>
__global__ __launch_bounds__(64 * W * (TY + 1)) void k_yee_fused(const KGrid g, const KFz z,
                                                                 int kchunk, int ntile_y,
                                                                 int nchunk, int n_out,
                                                                 int n_in, int seg, int order,
                                                                 TileBox bx) {
  __shared__ FusedLds<T, VEC, W, TY, XS> lds;
  // dispersive instances: the box table copied to LDS once per workgroup,
  // so the per-row scans read LDS instead of issuing global loads whose
  // waits would also wait for the plane prefetch in flight
  __shared__ KAde ade_tab[ADE ? kMaxAde : 1];
  // and the nine update coefficients of each box in the field type (the
  // E update then reads them without a conversion)
  __shared__ T ade_cf[ADE ? kMaxAde : 1][12];
  if constexpr (ADE) {
    static_assert(sizeof(KAde) % 4 == 0, "KAde is copied as 32-bit words");
    const int n = g.n_ade * (int)(sizeof(KAde) / 4);
    const gptr<const uint32_t> src = (gptr<const uint32_t>)g.ade;
    uint32_t* dst = (uint32_t*)ade_tab;
    for (int t = threadIdx.x; t < n; t += blockDim.x) dst[t] = src[t];
    __syncthreads();
    if ((int)threadIdx.x < g.n_ade * 9) {
      const int bq = threadIdx.x / 9, q = threadIdx.x % 9;
      const KAde& a = ade_tab[bq];
      const double v = q == 0 ? a.ke : q == 1 ? a.kp : q == 2 ? a.kq : q == 3 ? a.ks
                     : q == 4 ? a.a0 : q == 5 ? a.a1 : q == 6 ? a.a2 : q == 7 ? a.a3 : a.a4;
      ade_cf[bq][q] = (T)v;
    }
    __syncthreads();
  }
  const int b = blockIdx.x;
  if (b < seg)
    fused_tile<T, CK, VEC, W, TY, XS, true, ADE>(g, z, kchunk, ntile_y, nchunk, n_out, order,
                                                 bx, b, lds, ade_tab, ade_cf);
  else
    fused_tile<T, CK, VEC, W, TY, XS, false, ADE>(g, z, kchunk, ntile_y, nchunk, n_in, order,
                                                  bx, b - seg, lds, ade_tab, ade_cf);
}

template <typename T, int VEC, int W, int TY>
constexpr bool fused_fits() {
  return 64 * W * (TY + 1) <= 1024 && (size_t)9 * (TY + 1) * 64 * VEC * W * sizeof(T) <= 65536;
}

template <typename T, int CK, int VEC, int W, int TY, int XS>
int launch_fused_one(const KGrid& g, const KFz& z, int kchunk, int order, hipStream_t s) {
  if constexpr (fused_fits<T, VEC, W, TY>()) {
    const int ntile_y = (g.ny + TY - 1) / TY;
    const int nchunk = (g.nz + kchunk - 1) / kchunk;
    // interior box: a tile computes H on rows j0 .. j0+TY and E on rows
    // j0 .. j0+TY-1, planes kb .. ke (H) and kb .. ke-1 (E)
    TileBox bx{0, ntile_y, 0, nchunk};
    if (g.has_pml) {
      const int ylo = g.pml[2], yhi = g.ny - g.pml[3];  // rows [ylo, yhi) are inside
      const int zlo = g.pml[4], zhi = g.nz - g.pml[5];
      bx.ty0 = (ylo + TY - 1) / TY;
      bx.ty1 = std::max(bx.ty0, (yhi - 1 - TY) / TY + 1);
      if ((bx.ty1 - 1) * TY + TY > yhi - 1) bx.ty1 = bx.ty0;
      bx.kc0 = (zlo + kchunk - 1) / kchunk;
      bx.kc1 = std::max(bx.kc0, (zhi - 1) / kchunk);
      if (bx.kc1 > bx.kc0 && (bx.kc1 - 1) * kchunk + kchunk > zhi - 1) bx.kc1 = bx.kc0;
      if (bx.ty1 <= bx.ty0 || bx.kc1 <= bx.kc0) bx = TileBox{0, 0, 0, 0};
    }
    // XS workgroups (row segments) per tile
    const int n_in = (bx.ty1 - bx.ty0) * (bx.kc1 - bx.kc0) * XS;
    const int n_out = ntile_y * nchunk * XS - n_in;
    const int seg = ((n_out + 7) / 8) * 8;
    const int nwg = seg + ((n_in + 7) / 8) * 8;
    // plans with dispersive boxes take an instance with the ADE update
    // inlined (rows that cut a box, checked per box); the others carry no
    // ADE code at all
    if (g.n_ade)
      hipLaunchKernelGGL((k_yee_fused<T, CK, VEC, W, TY, XS, true>), dim3(nwg),
                         dim3(64 * W * (TY + 1)), 0, s, g, z, kchunk, ntile_y, nchunk, n_out,
                         n_in, seg, order, bx);
    else
      hipLaunchKernelGGL((k_yee_fused<T, CK, VEC, W, TY, XS, false>), dim3(nwg),
                         dim3(64 * W * (TY + 1)), 0, s, g, z, kchunk, ntile_y, nchunk, n_out,
                         n_in, seg, order, bx);
    return 0;
  } else {
    return -1;
  }
}

template <typename T, int CK, int VEC, int W, int XS = 1>
int launch_fused_w(const KGrid& g, const KFz& z, int ty, int kchunk, int order,
                   hipStream_t s) {
  if (ty == 2) return launch_fused_one<T, CK, VEC, W, 2, XS>(g, z, kchunk, order, s);
  if (ty == 3) return launch_fused_one<T, CK, VEC, W, 3, XS>(g, z, kchunk, order, s);
  if (ty == 4) return launch_fused_one<T, CK, VEC, W, 4, XS>(g, z, kchunk, order, s);
  // taller tiles for 128-cell rows or segments (up to 16 waves)
  if constexpr (VEC == 1 && W == 2) {
    if (ty == 5) return launch_fused_one<T, CK, VEC, W, 5, XS>(g, z, kchunk, order, s);
    if (ty == 7) return launch_fused_one<T, CK, VEC, W, 7, XS>(g, z, kchunk, order, s);
  }
  return -1;
}

// Shapes built with rows in two segments (one cell per lane in segments of
// 128 or 256 cells, or, FP32 only, two cells per lane in segments of 256;
// not with per-cell Ca/Da/Db).
constexpr bool fused_split_built(int esz, int ck, int vec, int w) {
  return ck != PM_COEFF_FULL &&
         ((vec == 1 && (w == 2 || w == 4)) || (esz == 4 && vec == 2 && w == 2));
}

// (vec, w) pairs with 64*vec*w = row width 128, 256 or 512 (xs = 1), or
// half of it (xs = 2)
template <typename T, int CK>
int launch_fused_ck(const KGrid& g, const KFz& z, int vec, int w, int ty, int kchunk,
                    int order, int xs, hipStream_t s) {
  if (xs == 2) {
    if constexpr (CK != PM_COEFF_FULL) {
      if (vec == 1 && w == 2) return launch_fused_w<T, CK, 1, 2, 2>(g, z, ty, kchunk, order, s);
      if (vec == 1 && w == 4) return launch_fused_w<T, CK, 1, 4, 2>(g, z, ty, kchunk, order, s);
      if constexpr (sizeof(T) == 4)
        if (vec == 2 && w == 2) return launch_fused_w<T, CK, 2, 2, 2>(g, z, ty, kchunk, order, s);
    }
    return -1;
  }
  if (xs != 1) return -1;
  if (vec == 1 && w == 2) return launch_fused_w<T, CK, 1, 2>(g, z, ty, kchunk, order, s);
  if (vec == 2 && w == 1) return launch_fused_w<T, CK, 2, 1>(g, z, ty, kchunk, order, s);
  if (vec == 1 && w == 4) return launch_fused_w<T, CK, 1, 4>(g, z, ty, kchunk, order, s);
  if (vec == 2 && w == 2) return launch_fused_w<T, CK, 2, 2>(g, z, ty, kchunk, order, s);
  if (vec == 4 && w == 1) return launch_fused_w<T, CK, 4, 1>(g, z, ty, kchunk, order, s);
  if (vec == 2 && w == 4) return launch_fused_w<T, CK, 2, 4>(g, z, ty, kchunk, order, s);
  if (vec == 4 && w == 2) return launch_fused_w<T, CK, 4, 2>(g, z, ty, kchunk, order, s);
  return -1;
}

// Offsets of the arrays from the lowest address of each group (field set
// read, field set written, everything else).  The kernel derives most
// offsets from a few strides, so it needs the packed layout the engine
// allocates: the six components of a field set at a constant stride, the
// three Cb arrays at a constant stride, the two CPML memory arrays of an
// axis back to back, the six profiles of an axis back to back.  False when
// the plan is laid out otherwise or a group spans 4 GiB (the caller then
// steps unfused).
inline bool make_kfz_impl(const KGrid& g, const FusedBufs& fb, int esz, KFz* out,
                          const char** why) {
  KFz z{};
  auto u = [](const void* p) { return (uintptr_t)p; };
#define PM_KFZ_FAIL(msg) \
  do {                   \
    *why = msg;          \
    return false;        \
  } while (0)
  const uint64_t field = (uint64_t)g.sz * (uint64_t)g.nz * (uint64_t)esz;
  const int n[3] = {g.nx, g.ny, g.nz};
  // field sets: Ex Ey Ez Hx Hy Hz at a constant stride
  const void* fin[6] = {fb.Ei[0], fb.Ei[1], fb.Ei[2], fb.Hi[0], fb.Hi[1], fb.Hi[2]};
  const void* fout[6] = {fb.Eo[0], fb.Eo[1], fb.Eo[2], fb.Ho[0], fb.Ho[1], fb.Ho[2]};
  for (int c = 0; c < 6; ++c)
    if (!fin[c] || !fout[c]) PM_KFZ_FAIL("missing field array");
  const intptr_t cs = (intptr_t)(u(fin[1]) - u(fin[0]));
  if (cs < (intptr_t)field || 6 * (uint64_t)cs >= (1ull << 32))
    PM_KFZ_FAIL("field set is not packed (component stride)");
  for (int c = 1; c < 6; ++c) {
    if ((intptr_t)(u(fin[c]) - u(fin[c - 1])) != cs) PM_KFZ_FAIL("input field set not packed");
    if ((intptr_t)(u(fout[c]) - u(fout[c - 1])) != cs) PM_KFZ_FAIL("output field set not packed");
  }
  z.bi = (const char*)fin[0];
  z.bo = (const char*)fout[0];
  z.cs = (uint32_t)cs;
  // everything else relative to the lowest of it
  const void* rest[64];
  int nr = 0;
  auto add = [&](const void* p) {
    if (p) rest[nr++] = p;
  };
  for (int a = 0; a < 3; ++a) {
    add(g.cb[a]);
    add(g.ca[a]);
    add(g.da[a]);
    add(g.db[a]);
    add(g.pml_be[a]);
    add(g.psi_e[a][0]);
    add(fb.psh_in[a][0]);
    add(fb.psh_out[a][0]);
    add(fb.pse_out[a][0]);
  }
  if (nr == 0) PM_KFZ_FAIL("no coefficient arrays");
  uintptr_t lo = ~(uintptr_t)0;
  for (int t = 0; t < nr; ++t) lo = std::min(lo, u(rest[t]));
  auto off = [&](const void* p, uint32_t* slot) {
    if (!p) {
      *slot = 0;
      return true;
    }
    const uint64_t d = (uint64_t)(u(p) - lo);
    if (d + 2 * field >= (1ull << 32)) return false;
    *slot = (uint32_t)d;
    return true;
  };
  z.bc = (const char*)lo;
  if (!g.cb[0] || !g.cb[1] || !g.cb[2]) PM_KFZ_FAIL("missing Cb");
  const intptr_t ccs = (intptr_t)(u(g.cb[1]) - u(g.cb[0]));
  if (ccs <= 0 || (intptr_t)(u(g.cb[2]) - u(g.cb[1])) != ccs) PM_KFZ_FAIL("Cb not packed");
  z.ccs = (uint32_t)ccs;
  if (!off(g.cb[0], &z.cb)) PM_KFZ_FAIL("Cb beyond 4 GiB");
  for (int a = 0; a < 3; ++a) {
    if (!off(g.ca[a], &z.ca[a]) || !off(g.da[a], &z.da[a]) || !off(g.db[a], &z.db[a]))
      PM_KFZ_FAIL("Ca/Da/Db beyond 4 GiB");
    if (g.pml[2 * a] + g.pml[2 * a + 1] == 0) {
      // no CPML on this axis: the kernel's unconditional CPML-memory loads
      // (slab position 0, values unused) read the Cb arrays instead, which
      // are larger than any position they can form
      z.pe[a] = z.phi[a] = z.pho[a] = z.prof[a] = z.peo[a] = z.cb;
      z.ps[a] = z.pn[a] = 0;
      continue;
    }
    // CPML memory pairs back to back, profiles back to back
    const uint64_t S = (uint64_t)(g.pml[2 * a] + g.pml[2 * a + 1]);
    const uint64_t ps = S * (uint64_t)g.nx * g.ny * g.nz / (uint64_t)n[a] * (uint64_t)esz;
    z.ps[a] = (uint32_t)ps;
    const void* pairs[3][2] = {{g.psi_e[a][0], g.psi_e[a][1]},
                               {fb.psh_in[a][0], fb.psh_in[a][1]},
                               {fb.psh_out[a][0], fb.psh_out[a][1]}};
    for (auto& pr : pairs)
      if (!pr[0] || !pr[1] || u(pr[1]) - u(pr[0]) != ps) PM_KFZ_FAIL("CPML memory pair not packed");
    if (!off(g.psi_e[a][0], &z.pe[a]) || !off(fb.psh_in[a][0], &z.phi[a]) ||
        !off(fb.psh_out[a][0], &z.pho[a]))
      PM_KFZ_FAIL("CPML memory beyond 4 GiB");
    z.peo[a] = z.pe[a];
    if (fb.pse_out[a][0]) {
      if (!fb.pse_out[a][1] || u(fb.pse_out[a][1]) - u(fb.pse_out[a][0]) != ps)
        PM_KFZ_FAIL("E-side CPML memory (written set) not packed");
      if (!off(fb.pse_out[a][0], &z.peo[a])) PM_KFZ_FAIL("CPML memory beyond 4 GiB");
    }
    const uintptr_t pn = (uintptr_t)n[a] * (uintptr_t)esz;
    const void* pr6[6] = {g.pml_be[a], g.pml_ce[a], g.pml_ike[a],
                          g.pml_bh[a], g.pml_ch[a], g.pml_ikh[a]};
    for (int t = 0; t < 6; ++t)
      if (!pr6[t] || u(pr6[t]) != u(pr6[0]) + (uintptr_t)t * pn) PM_KFZ_FAIL("profiles not packed");
    z.pn[a] = (uint32_t)pn;
    if (!off(g.pml_be[a], &z.prof[a])) PM_KFZ_FAIL("profiles beyond 4 GiB");
  }
  *out = z;
  return true;
#undef PM_KFZ_FAIL
}

// make_kfz_impl; PRISMO_AMD_DEBUG_FUSED=1 prints why a plan is not laid out
// for the fused kernel.
inline bool make_kfz(const KGrid& g, const FusedBufs& fb, int esz, KFz* out) {
  const char* why = "";
  const bool ok = make_kfz_impl(g, fb, esz, out, &why);
  if (!ok && std::getenv("PRISMO_AMD_DEBUG_FUSED")) std::fprintf(stderr, "[pmfdtd] fused layout: %s\n", why);
  return ok;
}

}  // namespace pm
