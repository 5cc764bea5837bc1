// Two Yee steps per pass (temporal blocking in the z-march) for FP32 fields
// with Cb in FP32 or FP16, CPML on any face, and the soft sources of the
// first of the two steps added inside the kernel.
//
// Arithmetic per cell and step is that of fused_kernel.h (textbook Yee,
// backward H differences, E beyond the grid = 0, Roden-Gedney CPML); what
// changes is how often a field element crosses HBM: each pass reads E(n),
// H(n-1/2) and the CPML memory once and writes E(n+2), H(n+3/2) once, so the
// HBM bytes per cell-step are half those of the one-step kernel.
//
// Decomposition.  A workgroup owns TY x-rows (j0 .. j0+TY-1) of a z-chunk
// [kb, ke) and works on a window of R = TY + 4 rows (j0-2 .. j0+TY+1), one
// row per W waves, VEC cells per lane.  The four half-steps of the pass lag
// each other by one plane in the z-march; iteration k:
//   H1 on plane k   (E(n) of planes k, k-1)                 -> H(n+1/2)
//   E1 on plane k-1 (H(n+1/2) of planes k-1, k)             -> E(n+1)
//   sources of step n on E(n+1) (plane k-1) and H(n+1/2) (plane k-1)
//   H2 on plane k-1 (E(n+1) of planes k-1, k-2)             -> H(n+3/2), stored
//   E2 on plane k-2 (H(n+3/2) of planes k-2, k-1)           -> E(n+2), stored
// Each half-step needs one row more on one side than the next, so H1 is
// valid on window rows 1..R-1, E1 on 1..R-2, H2 on 2..R-2 and E2 on the own
// rows 2..R-3; likewise the march runs over planes kb-2 .. ke+1.  Every row
// of the window computes everything (rows and planes outside the grid as
// zero), so the code has no row cases; values outside those cones are never
// used.  In-plane neighbours come from LDS: E of the current half-step (for
// the i-1 / j-1 differences of H) and H of the previous plane (for the i+1 /
// j+1 differences of E), three plane buffers, four barriers per iteration.
//
// Two kernels per pass.  Interior tiles (window inside the grid and outside
// the y and z CPML slabs) take TY = 4 rows and 16 waves; the rest (slab
// tiles) TY = 2 rows and 12 waves, which leaves them the registers for the
// y/z CPML terms.
//
// CPML memory.  A cell's step-1 memory (H side of plane k, E side of plane
// k-1) is used by step 2 one iteration later.  Interior tiles keep the x
// terms in LDS, one slot per lane and plane parity; slab tiles keep the x
// terms in registers and the y/z terms in one LDS slot per lane.  Both the
// H-side and the E-side memory are read from one buffer set and written to
// the other (rows and planes of the window are recomputed by up to three
// workgroups, so nothing is updated in place).
#pragma once
#include "fused_kernel.h"

namespace pm {

// One-step layout plus where the E-side CPML memory is written.
struct KFz2 {
  KFz z;             // z.pe: E-side memory read, z.peo: written
  int64_t src_row;   // table row of the first step of the pass
  int n_src;         // sources added inside the kernel (0: none)
  const KSrc* src;   // device table of the sources (the plan's scratch)
  // Lane map of a row taken by two waves: the second wave holds cells
  // [xa, xa + RW/2), the first the rest; xa > 0 when that puts every x CPML
  // cell in the first wave, so the second never runs the x CPML code
  // (0: lane l of wave w holds cells (64 w + l) VEC ..).
  int xa;
};

constexpr int kF2XSlab = 64;  // x CPML cells per row the LDS slots hold

template <int VEC, int W, int TY, bool YZ>
struct Fused2Lds {
  static constexpr int R = TY + 4, RW = 64 * VEC * W;
  float le[3][R][RW];        // E of the current half-step
  float lha[3][R][RW];       // H(n+1/2) of the previous plane
  float lhb[3][R][RW];       // H(n+3/2) of the previous plane
  float xp[6][RW];           // x profiles: be ce ike bh ch ikh
  // per-lane state that waits one iteration for step 2 (each lane reads
  // back what it wrote, so no barrier is involved):
  float pv[6][R][RW];        // E(n+1) and H(n+3/2) of the previous plane
  // step 1's x CPML memory by plane parity, indexed by slab cell: H 0 1, E 0 1
  float qx[2][4][R][kF2XSlab];
  // slab tiles: step 1's y/z CPML memory, H y0 y1 z0 z1 (plane k-1), E y0 y1
  // z0 z1 (plane k-2)
  float qyz[YZ ? 8 : 1][YZ ? R : 1][YZ ? RW : 1];
  KSrc src[kMaxSrc];
  double src_a[kMaxSrc];     // table value of each source at the first step
};

template <int VEC>
__device__ __forceinline__ void ldv(float (&d)[VEC], const float* p) {
  fld<float, VEC>(d, p);
}

// Sources of the first step of the pass on one row of one plane: E
// components into e[3], H components into h[3] (the reference adds every
// source after the E update of its step; FP32 fields add the value rounded
// to FP32, as the one-step source kernel does).  Only sources with one value
// per step (kind 0) come here, so the loop touches LDS alone.
template <int VEC, int W, int TY, bool YZ>
__device__ __forceinline__ void add_sources(const Fused2Lds<VEC, W, TY, YZ>& L, int n_src,
                                            int x0, int j, int kq, float (&e)[3][VEC],
                                            float (&h)[3][VEC]) {
  for (int s = 0; s < n_src; ++s) {
    const KSrc& q = L.src[s];
    if (j < q.j0 || j >= q.j1 || kq < q.k0 || kq >= q.k1) continue;
    const float fv = (float)L.src_a[s];
#pragma unroll
    for (int t = 0; t < VEC; ++t) {
      const int i = x0 + t;
      if (i < q.i0 || i >= q.i1) continue;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (q.comp == c) e[c][t] = e[c][t] + fv;
        if (q.comp == 3 + c) h[c][t] = h[c][t] + fv;
      }
    }
  }
}

// One tile: own rows j0 .. j0+TY-1 of planes [kb, ke).
template <int CK, int VEC, int W, int TY, bool YZ>
__device__ __forceinline__ void fused2_tile(const KGrid& g, const KFz2& z2, int j0, int kb,
                                            int ke, Fused2Lds<VEC, W, TY, YZ>& L) {
  using T = float;
  constexpr int R = TY + 4, RW = 64 * VEC * W;
  const KFz& z = z2.z;
  const int tid = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(tid / (64 * W));  // window row
  const int lane = tid % (64 * W);
  // wave of the row (uniform) and the first cell of this lane
  const int wx = __builtin_amdgcn_readfirstlane(lane / 64);
  int x0 = lane * VEC;
  bool xwave = true;  // this wave may hold x CPML cells
  if (W == 2 && z2.xa > 0) {
    const int l = (lane % 64) * VEC;
    x0 = wx == 1 ? z2.xa + l : (l < z2.xa ? l : l + RW / 2);
    xwave = wx == 0;
  }
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int j = j0 - 2 + w;
  // interior tiles: every window row and plane lies inside the grid
  const bool row_ok = !YZ || (j >= 0 && j < ny);
  const bool own = w >= 2 && w < TY + 2 && row_ok && j < ny;
  const int jr = YZ ? min(max(j, 0), ny - 1) : j;
  const T idx = (T)g.idx, idy = (T)g.idy, idz = (T)g.idz;
  const T db = (T)g.db_const;
  const uint32_t sy = (uint32_t)g.sy, sz = (uint32_t)g.sz;
  const uint32_t rowoff = (uint32_t)jr * sy + (uint32_t)x0;
  const int wm = w > 0 ? w - 1 : 0, wp = w + 1 < R ? w + 1 : R - 1;
  const bool pml = g.has_pml;

  // CPML: x per lane (profiles in LDS), y per row (scalars), z per plane
  int sxs[VEC];
  bool lane_pml_x = false;
#pragma unroll
  for (int e = 0; e < VEC; ++e) {
    const int i = x0 + e;
    sxs[e] = (pml && i < nx) ? fslab(i, nx, g.pml[0], g.pml[1]) : -1;
    lane_pml_x |= sxs[e] >= 0;
  }
  int sxc[VEC];
#pragma unroll
  for (int e = 0; e < VEC; ++e) sxc[e] = max(sxs[e], 0);
  const int sys = (YZ && pml && row_ok) ? fslab(j, ny, g.pml[2], g.pml[3]) : -1;
  const int syc = max(sys, 0);
  T ybh = 1, ych = 0, yikh = 1, ybe = 1, yce = 0, yike = 1;
  if (sys >= 0) {
    ybh = *zp<T>(z.bc, z.prof[1] + 3u * z.pn[1], j);
    ych = *zp<T>(z.bc, z.prof[1] + 4u * z.pn[1], j);
    yikh = *zp<T>(z.bc, z.prof[1] + 5u * z.pn[1], j);
    ybe = *zp<T>(z.bc, z.prof[1] + 0u * z.pn[1], j);
    yce = *zp<T>(z.bc, z.prof[1] + 1u * z.pn[1], j);
    yike = *zp<T>(z.bc, z.prof[1] + 2u * z.pn[1], j);
  }
  auto zprof = [&](int kq, int t) __attribute__((always_inline)) -> T {
    return *zp<T>(z.bc, z.prof[2] + (uint32_t)t * z.pn[2], kq);
  };

  // CPML memory of one plane: x pair (registers), y and z pairs (slab tiles)
  auto load_psi = [&](const uint32_t(&base)[3], int kq, T(&q)[6][VEC])
                      __attribute__((always_inline)) {
    const int kc_ = min(max(kq, 0), nz - 1);
    const int skc = YZ ? max(fslab(kc_, nz, g.pml[4], g.pml[5]), 0) : 0;
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      const int ic = min(x0 + e, nx - 1);
      const uint32_t ox = fslab_off(g, 0, sxc[e], ic, jr, kc_);
      q[0][e] = *zp<T>(z.bc, base[0], ox);
      q[1][e] = *zp<T>(z.bc, base[0] + z.ps[0], ox);
      if constexpr (YZ) {
        const uint32_t oy = fslab_off(g, 1, syc, ic, jr, kc_);
        q[2][e] = *zp<T>(z.bc, base[1], oy);
        q[3][e] = *zp<T>(z.bc, base[1] + z.ps[1], oy);
        const uint32_t oz = fslab_off(g, 2, skc, ic, jr, kc_);
        q[4][e] = *zp<T>(z.bc, base[2], oz);
        q[5][e] = *zp<T>(z.bc, base[2] + z.ps[2], oz);
      } else {
#pragma unroll
        for (int t = 2; t < 6; ++t) q[t][e] = (T)0;
      }
    }
  };

  struct PlaneIn {
    T e[3][VEC];   // E(n)
    T h[3][VEC];   // H(n-1/2)
  };
  auto load_plane = [&](int kq, PlaneIn& P) __attribute__((always_inline)) {
    const int kc_ = min(max(kq, 0), nz - 1);
    const uint32_t o = rowoff + (uint32_t)kc_ * sz;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      ldv<VEC>(P.e[c], zp<T>(z.bi, (uint32_t)c * z.cs, o));
      ldv<VEC>(P.h[c], zp<T>(z.bi, (uint32_t)(3 + c) * z.cs, o));
    }
  };

  // state carried from plane to plane
  T en_m[3][VEC], ha_m[3][VEC];
  CbRaw<T, CK> cb_m[3][VEC];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      en_m[c][e] = ha_m[c][e] = (T)0;
      cb_m[c][e] = (CbRaw<T, CK>)0;
    }
  T zh_m[3] = {1, 0, 1}, ze_m[3] = {1, 0, 1};  // z profiles of the planes step 2 works on
  const int n_src = z2.n_src;

  auto step = [&](const int k, const PlaneIn& cur, PlaneIn& nxt) __attribute__((always_inline)) {
    const int kk = k - 1, k2 = k - 2;
    const int kq = min(max(kk, 0), nz - 1);
    // (1) E-side inputs of E1 on plane k-1, ahead of the prefetch
    CbRaw<T, CK> cb[3][VEC];
    const uint32_t oq = rowoff + (uint32_t)kq * sz;
    fcb<T, CK, VEC>(z, 0, oq, cb[0]);
    fcb<T, CK, VEC>(z, 1, oq, cb[1]);
    fcb<T, CK, VEC>(z, 2, oq, cb[2]);
    T qe[6][VEC];
    load_psi(z.pe, kk, qe);
    T qh[6][VEC];   // H-side CPML memory (n-1/2) of plane k
    load_psi(z.phi, k, qh);
    T zh[3] = {1, 0, 1}, ze[3] = {1, 0, 1};
    if constexpr (YZ) {
      const int kh_ = min(max(k, 0), nz - 1);
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        zh[t] = zprof(kh_, 3 + t);
        ze[t] = zprof(kq, t);
      }
    }
    // (2) prefetch plane k+1
    load_plane(k + 1, nxt);
    // x profiles and slab slots are re-read from LDS each plane (an index
    // the compiler cannot prove loop-invariant), not held in registers
    int xo = x0;
    asm volatile("" : "+v"(xo));
    const bool pk = !YZ || (row_ok && k >= 0 && k < nz);
    const bool pkk = !YZ || (row_ok && kk >= 0 && kk < nz);
    const bool pk2 = !YZ || (row_ok && k2 >= 0 && k2 < nz);
    const int szk = (YZ && pml && pk) ? fslab(k, nz, g.pml[4], g.pml[5]) : -1;
    const int szkk = (YZ && pml && pkk) ? fslab(kk, nz, g.pml[4], g.pml[5]) : -1;
    const int szk2 = (YZ && pml && pk2) ? fslab(k2, nz, g.pml[4], g.pml[5]) : -1;
    // (3) share E(n) of plane k
    T en[3][VEC];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int e = 0; e < VEC; ++e) {
        en[c][e] = pk ? cur.e[c][e] : (T)0;
        L.le[c][w][x0 + e] = en[c][e];
      }
    __syncthreads();
    // (4) H1 on plane k, E1 on plane k-1
    T ha[3][VEC], ea[3][VEC];
    T qhy[4][VEC], qey[4][VEC];          // y/z CPML memory written by step 1
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      const int i = x0 + e;
      const T ex = en[0][e], ey = en[1][e], ez = en[2][e];
      const T exj = L.le[0][wm][i], ezj = L.le[2][wm][i];
      const T ezi = i > 0 ? L.le[2][w][i - 1] : (T)0;
      const T eyi = i > 0 ? L.le[1][w][i - 1] : (T)0;
      T c1 = (ez - ezj) * idy;  // Hx: dEz/dy - dEy/dz
      T c2 = (ey - en_m[1][e]) * idz;
      T d1 = (ex - en_m[0][e]) * idz;  // Hy: dEx/dz - dEz/dx
      T d2 = (ez - ezi) * idx;
      T f1 = (ey - eyi) * idx;  // Hz: dEy/dx - dEx/dy
      T f2 = (ex - exj) * idy;
#pragma unroll
      for (int t = 0; t < 4; ++t) qhy[t][e] = (T)0;
      if (pml) {
        if (xwave && sxs[e] >= 0) {
          const T bh = L.xp[3][xo + e], ch = L.xp[4][xo + e], ikh = L.xp[5][xo + e];
          const T p0 = bh * qh[0][e] + ch * d2, p1 = bh * qh[1][e] + ch * f1;
          L.qx[k & 1][0][w][sxs[e]] = p0;
          L.qx[k & 1][1][w][sxs[e]] = p1;
          d2 = d2 * ikh + p0;
          f1 = f1 * ikh + p1;
        }
        if constexpr (YZ) {
          if (sys >= 0) {
            const T p1 = ybh * qh[3][e] + ych * c1, p0 = ybh * qh[2][e] + ych * f2;
            qhy[0][e] = p0;
            qhy[1][e] = p1;
            c1 = c1 * yikh + p1;
            f2 = f2 * yikh + p0;
          }
          if (szk >= 0) {
            const T p0 = zh[0] * qh[4][e] + zh[1] * c2, p1 = zh[0] * qh[5][e] + zh[1] * d1;
            qhy[2][e] = p0;
            qhy[3][e] = p1;
            c2 = c2 * zh[2] + p0;
            d1 = d1 * zh[2] + p1;
          }
        }
      }
      T hx = cur.h[0][e] - db * (c1 - c2);
      T hy = cur.h[1][e] - db * (d1 - d2);
      T hz = cur.h[2][e] - db * (f1 - f2);
      if (!pk || i >= nx - 1) hx = 0;
      if (!pk || i >= nx || j >= ny - 1) hy = 0;
      if (!pk || i >= nx || k >= nz - 1) hz = 0;
      ha[0][e] = hx;
      ha[1][e] = hy;
      ha[2][e] = hz;
    }
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      const int i = x0 + e;
      const int ir = i + 1 < RW ? i + 1 : i;
      const T hyx = L.lha[1][w][ir], hzx = L.lha[2][w][ir];
      const T hxy = L.lha[0][wp][i], hzy = L.lha[2][wp][i];
      T a1 = (hzy - ha_m[2][e]) * idy;  // Ex: dHz/dy - dHy/dz
      T a2 = (ha[1][e] - ha_m[1][e]) * idz;
      T b1 = (ha[0][e] - ha_m[0][e]) * idz;  // Ey: dHx/dz - dHz/dx
      T b2 = (hzx - ha_m[2][e]) * idx;
      T c1 = (hyx - ha_m[1][e]) * idx;  // Ez: dHy/dx - dHx/dy
      T c2 = (hxy - ha_m[0][e]) * idy;
#pragma unroll
      for (int t = 0; t < 4; ++t) qey[t][e] = (T)0;
      if (pml) {
        if (xwave && sxs[e] >= 0) {
          const T be = L.xp[0][xo + e], ce = L.xp[1][xo + e], ike = L.xp[2][xo + e];
          const T p0 = be * qe[0][e] + ce * b2, p1 = be * qe[1][e] + ce * c1;
          L.qx[k & 1][2][w][sxs[e]] = p0;
          L.qx[k & 1][3][w][sxs[e]] = p1;
          b2 = b2 * ike + p0;
          c1 = c1 * ike + p1;
        }
        if constexpr (YZ) {
          if (sys >= 0) {
            const T p1 = ybe * qe[3][e] + yce * a1, p0 = ybe * qe[2][e] + yce * c2;
            qey[0][e] = p0;
            qey[1][e] = p1;
            a1 = a1 * yike + p1;
            c2 = c2 * yike + p0;
          }
          if (szkk >= 0) {
            const T p0 = ze[0] * qe[4][e] + ze[1] * a2, p1 = ze[0] * qe[5][e] + ze[1] * b1;
            qey[2][e] = p0;
            qey[3][e] = p1;
            a2 = a2 * ze[2] + p0;
            b1 = b1 * ze[2] + p1;
          }
        }
      }
      T vx = en_m[0][e] + fcbv<T, CK>(g, cb[0][e]) * (a1 - a2);
      T vy = en_m[1][e] + fcbv<T, CK>(g, cb[1][e]) * (b1 - b2);
      T vz = en_m[2][e] + fcbv<T, CK>(g, cb[2][e]) * (c1 - c2);
      if (!pkk || i >= nx || j >= ny - 1 || kk >= nz - 1) vx = 0;
      if (!pkk || i >= nx - 1 || kk >= nz - 1) vy = 0;
      if (!pkk || i >= nx - 1 || j >= ny - 1) vz = 0;
      ea[0][e] = vx;
      ea[1][e] = vy;
      ea[2][e] = vz;
    }
    // sources of the first step: E(n+1) and H(n+1/2) of plane k-1 (after E1
    // used H(n+1/2), before step 2 uses either)
    if (n_src > 0 && pkk) add_sources<VEC, W, TY, YZ>(L, n_src, x0, j, kk, ea, ha_m);
    __syncthreads();
    // (5) share H(n+1/2) of plane k and E(n+1) of plane k-1
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int e = 0; e < VEC; ++e) {
        L.lha[c][w][x0 + e] = ha[c][e];
        L.le[c][w][x0 + e] = ea[c][e];
      }
    __syncthreads();
    // (6) H2 on plane k-1 from E(n+1) of planes k-1 (LDS) and k-2
    T hb[3][VEC];
    const bool store_h = own && kk >= kb && kk < ke;
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      const int i = x0 + e;
      const T ex = ea[0][e], ey = ea[1][e], ez = ea[2][e];
      const T exj = L.le[0][wm][i], ezj = L.le[2][wm][i];
      const T ezi = i > 0 ? L.le[2][w][i - 1] : (T)0;
      const T eyi = i > 0 ? L.le[1][w][i - 1] : (T)0;
      T c1 = (ez - ezj) * idy;
      T c2 = (ey - L.pv[1][w][i]) * idz;   // E(n+1) of plane k-2
      T d1 = (ex - L.pv[0][w][i]) * idz;
      T d2 = (ez - ezi) * idx;
      T f1 = (ey - eyi) * idx;
      T f2 = (ex - exj) * idy;
      if (pml && i < nx) {
        if (xwave && sxs[e] >= 0) {
          const uint32_t o = fslab_off(g, 0, sxs[e], i, j, kk);
          const T bh = L.xp[3][xo + e], ch = L.xp[4][xo + e], ikh = L.xp[5][xo + e];
          const T q0 = L.qx[(k - 1) & 1][0][w][sxs[e]];
          const T q1 = L.qx[(k - 1) & 1][1][w][sxs[e]];
          const T p0 = bh * q0 + ch * d2, p1 = bh * q1 + ch * f1;
          if (store_h && j < ny - 1) *zpw<T>(z.bc, z.pho[0], o) = p0;
          if (store_h && kk < nz - 1) *zpw<T>(z.bc, z.pho[0] + z.ps[0], o) = p1;
          d2 = d2 * ikh + p0;
          f1 = f1 * ikh + p1;
        }
        if constexpr (YZ) {
          const int li = lane * VEC + e;
          if (sys >= 0) {
            const uint32_t o = fslab_off(g, 1, sys, i, j, kk);
            const T p0 = ybh * L.qyz[0][w][li] + ych * f2;
            const T p1 = ybh * L.qyz[1][w][li] + ych * c1;
            if (store_h && i < nx - 1) *zpw<T>(z.bc, z.pho[1] + z.ps[1], o) = p1;
            if (store_h && kk < nz - 1) *zpw<T>(z.bc, z.pho[1], o) = p0;
            c1 = c1 * yikh + p1;
            f2 = f2 * yikh + p0;
          }
          if (szkk >= 0) {
            const uint32_t o = fslab_off(g, 2, szkk, i, j, kk);
            const T p0 = zh_m[0] * L.qyz[2][w][li] + zh_m[1] * c2;
            const T p1 = zh_m[0] * L.qyz[3][w][li] + zh_m[1] * d1;
            if (store_h && i < nx - 1) *zpw<T>(z.bc, z.pho[2], o) = p0;
            if (store_h && j < ny - 1) *zpw<T>(z.bc, z.pho[2] + z.ps[2], o) = p1;
            c2 = c2 * zh_m[2] + p0;
            d1 = d1 * zh_m[2] + p1;
          }
        }
      }
      T hx = ha_m[0][e] - db * (c1 - c2);
      T hy = ha_m[1][e] - db * (d1 - d2);
      T hz = ha_m[2][e] - db * (f1 - f2);
      if (!pkk || i >= nx - 1) hx = 0;
      if (!pkk || i >= nx || j >= ny - 1) hy = 0;
      if (!pkk || i >= nx || kk >= nz - 1) hz = 0;
      hb[0][e] = hx;
      hb[1][e] = hy;
      hb[2][e] = hz;
    }
    if (store_h) {
      const uint32_t o = rowoff + (uint32_t)kk * sz;
      fst<T, VEC>(zpw<T>(z.bo, 3u * z.cs, o), hb[0]);
      fst<T, VEC>(zpw<T>(z.bo, 4u * z.cs, o), hb[1]);
      fst<T, VEC>(zpw<T>(z.bo, 5u * z.cs, o), hb[2]);
    }
    // (7) E2 on plane k-2 from H(n+3/2) of planes k-2 (LDS) and k-1
    if (own && k2 >= kb && k2 < ke) {
      T ox[VEC], oy[VEC], oz[VEC];
#pragma unroll
      for (int e = 0; e < VEC; ++e) {
        const int i = x0 + e;
        const int ir = i + 1 < RW ? i + 1 : i;
        const T hyx = L.lhb[1][w][ir], hzx = L.lhb[2][w][ir];
        const T hxy = L.lhb[0][wp][i], hzy = L.lhb[2][wp][i];
        // H(n+3/2) of plane k-2 from this lane's slots
        const T hbx = L.pv[3][w][i], hby = L.pv[4][w][i], hbz = L.pv[5][w][i];
        T a1 = (hzy - hbz) * idy;
        T a2 = (hb[1][e] - hby) * idz;
        T b1 = (hb[0][e] - hbx) * idz;
        T b2 = (hzx - hbz) * idx;
        T c1 = (hyx - hby) * idx;
        T c2 = (hxy - hbx) * idy;
        const bool exok = i < nx && j < ny - 1 && k2 < nz - 1;
        const bool eyok = i < nx - 1 && k2 < nz - 1;
        const bool ezok = i < nx - 1 && j < ny - 1;
        if (pml) {
          if (xwave && sxs[e] >= 0) {
            const uint32_t o = fslab_off(g, 0, sxs[e], i, j, k2);
            const T be = L.xp[0][xo + e], ce = L.xp[1][xo + e], ike = L.xp[2][xo + e];
            const T q0 = L.qx[(k - 1) & 1][2][w][sxs[e]];
            const T q1 = L.qx[(k - 1) & 1][3][w][sxs[e]];
            const T p0 = be * q0 + ce * b2, p1 = be * q1 + ce * c1;
            if (eyok) *zpw<T>(z.bc, z.peo[0], o) = p0;
            if (ezok) *zpw<T>(z.bc, z.peo[0] + z.ps[0], o) = p1;
            b2 = b2 * ike + p0;
            c1 = c1 * ike + p1;
          }
          if constexpr (YZ) {
            const int li = lane * VEC + e;
            if (sys >= 0) {
              const uint32_t o = fslab_off(g, 1, sys, i, j, k2);
              const T p0 = ybe * L.qyz[4][w][li] + yce * c2;
              const T p1 = ybe * L.qyz[5][w][li] + yce * a1;
              if (exok) *zpw<T>(z.bc, z.peo[1] + z.ps[1], o) = p1;
              if (ezok) *zpw<T>(z.bc, z.peo[1], o) = p0;
              a1 = a1 * yike + p1;
              c2 = c2 * yike + p0;
            }
            if (szk2 >= 0) {
              const uint32_t o = fslab_off(g, 2, szk2, i, j, k2);
              const T p0 = ze_m[0] * L.qyz[6][w][li] + ze_m[1] * a2;
              const T p1 = ze_m[0] * L.qyz[7][w][li] + ze_m[1] * b1;
              if (exok) *zpw<T>(z.bc, z.peo[2], o) = p0;
              if (eyok) *zpw<T>(z.bc, z.peo[2] + z.ps[2], o) = p1;
              a2 = a2 * ze_m[2] + p0;
              b1 = b1 * ze_m[2] + p1;
            }
          }
        }
        const T vx = L.pv[0][w][i] + fcbv<T, CK>(g, cb_m[0][e]) * (a1 - a2);
        const T vy = L.pv[1][w][i] + fcbv<T, CK>(g, cb_m[1][e]) * (b1 - b2);
        const T vz = L.pv[2][w][i] + fcbv<T, CK>(g, cb_m[2][e]) * (c1 - c2);
        ox[e] = exok ? vx : (T)0;
        oy[e] = eyok ? vy : (T)0;
        oz[e] = ezok ? vz : (T)0;
      }
      const uint32_t o = rowoff + (uint32_t)k2 * sz;
      fst<T, VEC>(zpw<T>(z.bo, 0u, o), ox);
      fst<T, VEC>(zpw<T>(z.bo, z.cs, o), oy);
      fst<T, VEC>(zpw<T>(z.bo, 2u * z.cs, o), oz);
    }
    __syncthreads();
    // (8) share H(n+3/2) of plane k-1; step 1's y/z CPML memory waits in
    //     LDS for step 2 (each lane its own slots: no barrier needed)
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int e = 0; e < VEC; ++e) L.lhb[c][w][x0 + e] = hb[c][e];
    if constexpr (YZ) {
#pragma unroll
      for (int e = 0; e < VEC; ++e) {
        const int li = lane * VEC + e;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          L.qyz[t][w][li] = qhy[t][e];
          L.qyz[4 + t][w][li] = qey[t][e];
        }
      }
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        zh_m[t] = zh[t];
        ze_m[t] = ze[t];
      }
    }
    // (9) plane k becomes plane k-1
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int e = 0; e < VEC; ++e) {
        en_m[c][e] = en[c][e];
        ha_m[c][e] = ha[c][e];
        L.pv[c][w][x0 + e] = ea[c][e];
        L.pv[3 + c][w][x0 + e] = hb[c][e];
        cb_m[c][e] = cb[c][e];
      }
  };

  PlaneIn pa, pb;
  load_plane(kb - 2, pa);
  for (int k = kb - 2; k <= ke + 1; k += 2) {
    step(k, pa, pb);
    if (k + 1 > ke + 1) break;
    step(k + 1, pb, pa);
  }
}

// Workgroup b -> tile index: order 2 keeps the hardware's round-robin over
// the XCDs, order 3 groups the tile list by 64 so each XCD takes runs of 8
// neighbouring tiles (their shared halo rows become L2 hits), order 0 gives
// each XCD a contiguous range.
__device__ __forceinline__ int fused2_order(int b, int nblocks, int order) {
  if (order == 2) return b;
  if (order == 3) return b < (nblocks & ~63) ? (b & ~63) + (b & 7) * 8 + ((b >> 3) & 7) : b;
  const int per = (nblocks + 7) / 8;
  return (b % 8) * per + b / 8;
}

// Shared set-up of both kernels: x profiles (b = 1, c = 0, 1/kappa = 1
// outside the layers) and the sources of the first step into LDS.
template <int VEC, int W, int TY, bool YZ>
__device__ __forceinline__ void fused2_setup(const KGrid& g, const KFz2& z2,
                                             Fused2Lds<VEC, W, TY, YZ>& L) {
  constexpr int RW = 64 * VEC * W;
  const KFz& z = z2.z;
  for (int t = threadIdx.x; t < 6 * RW; t += blockDim.x) {
    const int r = t / RW, i = t % RW;
    float v = (r == 1 || r == 4) ? 0.f : 1.f;
    if (g.has_pml && i < g.nx && fslab(i, g.nx, g.pml[0], g.pml[1]) >= 0)
      v = *zp<float>(z.bc, z.prof[0] + (uint32_t)r * z.pn[0], i);
    L.xp[r][i] = v;
  }
  if ((int)threadIdx.x < z2.n_src) {
    L.src[threadIdx.x] = z2.src[threadIdx.x];
    L.src_a[threadIdx.x] = ((gptr<const double>)z2.src[threadIdx.x].table)[z2.src_row];
  }
  __syncthreads();
}

// Interior tiles: box [ty0, ty1) x [kc0, kc1) of TY-row tiles and z-chunks.
template <int CK, int VEC, int W, int TY>
__global__ __launch_bounds__(64 * W * (TY + 4)) void k_yee_fused2_in(const KGrid g,
                                                                      const KFz2 z2,
                                                                      int kchunk, TileBox bx,
                                                                      int nblocks, int order) {
  __shared__ Fused2Lds<VEC, W, TY, false> L;
  const int b = fused2_order(blockIdx.x, nblocks, order);
  if (b >= nblocks) return;
  fused2_setup<VEC, W, TY, false>(g, z2, L);
  const int nty = bx.ty1 - bx.ty0;
  const int ty = bx.ty0 + b % nty, kc = bx.kc0 + b / nty;
  const int kb = kc * kchunk;
  fused2_tile<CK, VEC, W, TY, false>(g, z2, ty * TY, kb, min(kb + kchunk, g.nz), L);
}

// Rows and chunks the slab kernel covers: every z-chunk outside [kc0, kc1)
// whole, and in those chunks the rows below y_lo and from y_hi on, in tiles
// of TY rows.
struct SlabMap {
  int kc0, kc1, nchunk;
  int y_lo, y_hi;         // rows [y_lo, y_hi) of chunks [kc0, kc1) are interior
  int per_full, lo_tiles, hi_tiles;
};

template <int CK, int VEC, int W, int TY>
__global__ __launch_bounds__(64 * W * (TY + 4)) void k_yee_fused2_slab(const KGrid g,
                                                                        const KFz2 z2,
                                                                        int kchunk, SlabMap m,
                                                                        int nblocks, int order) {
  __shared__ Fused2Lds<VEC, W, TY, true> L;
  const int b = fused2_order(blockIdx.x, nblocks, order);
  if (b >= nblocks) return;
  fused2_setup<VEC, W, TY, true>(g, z2, L);
  const int n_lo = m.kc0 * m.per_full, n_hi = (m.nchunk - m.kc1) * m.per_full;
  int kc, j0;
  if (b < n_lo) {
    kc = b / m.per_full;
    j0 = (b % m.per_full) * TY;
  } else if (b < n_lo + n_hi) {
    kc = m.kc1 + (b - n_lo) / m.per_full;
    j0 = ((b - n_lo) % m.per_full) * TY;
  } else {
    const int per = m.lo_tiles + m.hi_tiles;
    const int r = b - n_lo - n_hi;
    kc = m.kc0 + r / per;
    const int t = r % per;
    j0 = t < m.lo_tiles ? t * TY : m.y_hi + (t - m.lo_tiles) * TY;
  }
  const int kb = kc * kchunk;
  fused2_tile<CK, VEC, W, TY, true>(g, z2, j0, kb, min(kb + kchunk, g.nz), L);
}

constexpr int kF2TyIn = 4;    // rows per interior tile
constexpr int kF2TySlab = 2;  // rows per slab tile

template <int VEC, int W>
constexpr bool fused2_fits() {
  return 64 * W * (kF2TyIn + 4) <= 1024 && 64 * W * (kF2TySlab + 4) <= 1024 &&
         sizeof(Fused2Lds<VEC, W, kF2TyIn, false>) <= 160 * 1024 &&
         sizeof(Fused2Lds<VEC, W, kF2TySlab, true>) <= 160 * 1024;
}

// Interior box: tiles whose window rows [j0-2, j0+TY+1] and planes
// [kb-2, ke+1] lie inside the grid and outside the y and z CPML slabs.
inline TileBox fused2_box(const KGrid& g, int TY, int kchunk, int ntile_y, int nchunk) {
  const int ylo = g.has_pml ? g.pml[2] : 0, yhi = g.ny - (g.has_pml ? g.pml[3] : 0);
  const int zlo = g.has_pml ? g.pml[4] : 0, zhi = g.nz - (g.has_pml ? g.pml[5] : 0);
  int ty0 = ntile_y, ty1 = 0;
  for (int t = 0; t < ntile_y; ++t)
    if (t * TY - 2 >= ylo && t * TY + TY + 1 <= yhi - 1) {
      ty0 = std::min(ty0, t);
      ty1 = std::max(ty1, t + 1);
    }
  int kc0 = nchunk, kc1 = 0;
  for (int c = 0; c < nchunk; ++c) {
    const int kb = c * kchunk, ke = std::min(kb + kchunk, g.nz);
    if (kb - 2 >= zlo && ke + 1 <= zhi - 1) {
      kc0 = std::min(kc0, c);
      kc1 = std::max(kc1, c + 1);
    }
  }
  if (ty1 <= ty0 || kc1 <= kc0) return TileBox{0, 0, 0, 0};
  return TileBox{ty0, ty1, kc0, kc1};
}

struct F2Streams {
  hipStream_t s, s2;
  hipEvent_t fork, join;
};

template <int CK, int VEC, int W>
int launch_fused2_one(const KGrid& g, const KFz2& z2, int kchunk, int order,
                      const F2Streams& st) {
  const hipStream_t s = st.s;
  if constexpr (fused2_fits<VEC, W>()) {
    const int ntile_y = (g.ny + kF2TyIn - 1) / kF2TyIn;
    const int nchunk = (g.nz + kchunk - 1) / kchunk;
    const TileBox bx = fused2_box(g, kF2TyIn, kchunk, ntile_y, nchunk);
    const int n_in = (bx.ty1 - bx.ty0) * (bx.kc1 - bx.kc0);
    SlabMap m{};
    m.nchunk = nchunk;
    if (n_in > 0) {
      m.kc0 = bx.kc0;
      m.kc1 = bx.kc1;
      m.y_lo = bx.ty0 * kF2TyIn;
      m.y_hi = bx.ty1 * kF2TyIn;
    } else {
      m.kc0 = m.kc1 = 0;  // every chunk whole
      m.y_lo = m.y_hi = 0;
    }
    m.per_full = (g.ny + kF2TySlab - 1) / kF2TySlab;
    m.lo_tiles = m.y_lo / kF2TySlab;
    m.hi_tiles = (g.ny - m.y_hi + kF2TySlab - 1) / kF2TySlab;
    const int n_slab = (m.kc0 + (nchunk - m.kc1)) * m.per_full +
                       (m.kc1 - m.kc0) * (m.lo_tiles + m.hi_tiles);
    // both kinds read one buffer set and write the other on disjoint cells:
    // the slab tiles run on the second stream beside the interior tiles
    const bool fork = st.s2 && n_slab > 0 && n_in > 0;
    hipStream_t ss = s;
    if (fork) {
      if (hipEventRecord(st.fork, s) != hipSuccess || hipStreamWaitEvent(st.s2, st.fork, 0) != hipSuccess)
        return -2;
      ss = st.s2;
    }
    if (n_slab > 0)
      hipLaunchKernelGGL((k_yee_fused2_slab<CK, VEC, W, kF2TySlab>), dim3((n_slab + 7) / 8 * 8),
                         dim3(64 * W * (kF2TySlab + 4)), 0, ss, g, z2, kchunk, m, n_slab, order);
    if (n_in > 0)
      hipLaunchKernelGGL((k_yee_fused2_in<CK, VEC, W, kF2TyIn>), dim3((n_in + 7) / 8 * 8),
                         dim3(64 * W * (kF2TyIn + 4)), 0, s, g, z2, kchunk, bx, n_in, order);
    if (fork) {
      if (hipEventRecord(st.join, st.s2) != hipSuccess || hipStreamWaitEvent(s, st.join, 0) != hipSuccess)
        return -2;
    }
    return 0;
  } else {
    return -1;
  }
}

}  // namespace pm
