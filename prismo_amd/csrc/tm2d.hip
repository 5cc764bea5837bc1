// One launch per textbook Yee step on a 2-D TM grid (Ez, Hx, Hy) of one
// simulation: H(n+1/2) and E(n+1) in a single pass, reading field set A and
// writing set B (ping-pong), CPML in the x and y slabs.
//
// Arithmetic per cell is that of the unfused 2-D kernels in yee.hip (the
// H update takes backward differences of E, the E update forward
// differences of H, CPML per Roden-Gedney with the H-side memory of set A
// read and that of set B written, the E-side memory updated in place); what
// changes is that one wave owns a strip of 64 columns over a run of rows
// and marches them: H at row j, then E at row j-1 from the H rows j-1 and j
// it just computed.  x neighbours come from lane shuffles; lane 0 takes
// Ez(x0-1) from one extra load, and the Hy(n+1/2) at x0+64 that lane 63's E
// update needs is recomputed from that cell's inputs with the same
// arithmetic (not stored).  The tile's last H row (its neighbour's first)
// is computed and not stored, so tiles never see half-updated values.
#include "pm_internal.h"

namespace pm {
namespace {

constexpr int kWaves2d = 4;  // waves per workgroup

template <typename T>
__device__ __forceinline__ T ld(const void* base, uint32_t e) {
  return *((gptr<const T>)base + e);
}

template <typename T>
__device__ __forceinline__ void st(void* base, uint32_t e, T v) {
  __builtin_nontemporal_store(v, (gptr<T>)base + e);
}

template <typename T>
__device__ __forceinline__ T lane_val(T v, int l) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
  } else {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
  }
}

struct Tm2dArgs {
  const void* ez_a;
  const void* hx_a;
  const void* hy_a;
  void* ez_b;
  void* hx_b;
  void* hy_b;
  const void* psh_x_a;  // H-side CPML memory: x slabs (Hy), y slabs (Hx)
  const void* psh_y_a;
  void* psh_x_b;
  void* psh_y_b;
  void* pse_x;          // E-side CPML memory (Ez), x and y slabs, in place
  void* pse_y;
  const void* cb;       // Cb of Ez (field type, or FP16 times cb_scale)
  int chunk;            // rows per tile
};

template <typename T>
struct TmRowH {
  T ez, hx, hy, psx, psy;
  T halo;  // lane 0: Ez(x0-1); 1: Ez(x0+64); 2: Hy(x0+64); 3: its x CPML memory
};

template <typename T, int CK>
struct TmRowE {
  using CbT = typename std::conditional<CK == PM_COEFF_CB_F16, unsigned short, T>::type;
  CbT cb;
  T pex, pey;
};

template <typename T, int CK>
__global__ __launch_bounds__(64 * kWaves2d) void k_tm2d_fused(const KGrid g, const Tm2dArgs a,
                                                             int n_items) {
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int item = blockIdx.x * kWaves2d + wv;
  if (item >= n_items) return;  // whole wave
  const int nx = g.nx, ny = g.ny;
  const int nstrip = (nx + 63) / 64;
  const int strip = item % nstrip, tile = item / nstrip;
  const int x0 = strip * 64;
  const int i = x0 + lane;
  const bool inx = i < nx;
  const int ic = min(i, nx - 1);
  const uint32_t syu = (uint32_t)g.sy;
  const T idx = (T)g.idx, idy = (T)g.idy, db = (T)g.db_const, cbs = (T)g.cb_scale;
  // CPML: slab indices and profiles (b, c, 1/kappa; H side at nodes, E side
  // at half nodes), identity outside the slabs
  const int px0 = g.pml[0], px1 = g.pml[1], py0 = g.pml[2], py1 = g.pml[3];
  const int Sx = px0 + px1;
  auto slab = [](int p, int n, int lo, int hi) {
    return p < lo ? p : (p >= n - hi ? lo + (p - (n - hi)) : -1);
  };
  const int sxs = (g.has_pml && inx) ? slab(i, nx, px0, px1) : -1;
  const bool xpml = sxs >= 0;
  T bhx = 1, chx = 0, ikhx = 1, bex = 1, cex = 0, ikex = 1;
  if (xpml) {
    bhx = ((const T*)g.pml_bh[0])[i];
    chx = ((const T*)g.pml_ch[0])[i];
    ikhx = ((const T*)g.pml_ikh[0])[i];
    bex = ((const T*)g.pml_be[0])[i];
    cex = ((const T*)g.pml_ce[0])[i];
    ikex = ((const T*)g.pml_ike[0])[i];
  }
  // the cell right of the strip, x0+64 (uniform)
  const int ir = x0 + 64;
  const bool rin = ir < nx;
  const int irc = min(ir, nx - 1);
  const int sxr = (g.has_pml && rin) ? slab(ir, nx, px0, px1) : -1;
  const bool xpml_r = sxr >= 0;
  T bhx_r = 1, chx_r = 0, ikhx_r = 1;
  if (xpml_r) {
    bhx_r = ((const T*)g.pml_bh[0])[ir];
    chx_r = ((const T*)g.pml_ch[0])[ir];
    ikhx_r = ((const T*)g.pml_ikh[0])[ir];
  }
  // y profiles: read-only for the launch, scalar loads when a row needs them
  using CT = const T __attribute__((address_space(4)))*;
  const CT ybh = (CT)g.pml_bh[1], ych = (CT)g.pml_ch[1], yikh = (CT)g.pml_ikh[1];
  const CT ybe = (CT)g.pml_be[1], yce = (CT)g.pml_ce[1], yike = (CT)g.pml_ike[1];
  using CbT = typename TmRowE<T, CK>::CbT;

  // psi offsets: x slabs s + Sx j, y slabs i + nx s
  auto xoff = [&](int s, int j) { return (uint32_t)s + (uint32_t)Sx * (uint32_t)j; };
  auto yoff = [&](int ii, int s) { return (uint32_t)ii + (uint32_t)nx * (uint32_t)s; };

  // Every load of the row loop is issued on every path from an in-range
  // position (a value outside its use is masked), so the compiler's memory
  // counter waits never cover the rows prefetched.  The halo load: lane 0
  // Ez(x0-1) (element 0 for the first strip), lane 1 Ez(x0+64), lane 2
  // Hy(x0+64), lane 3 the Hy CPML memory there (element 0 outside the x
  // slabs), other lanes Ez of their own cell.
  const void* hbase = lane == 2 ? a.hy_a : (lane == 3 ? a.psh_x_a : a.ez_a);
  auto load_h = [&](int j, TmRowH<T>& r) __attribute__((always_inline)) {
    const int jc = min(j, ny - 1);
    const uint32_t o = (uint32_t)jc * syu + (uint32_t)ic;
    uint32_t ho;
    if (lane == 0) ho = (uint32_t)jc * syu + (uint32_t)max(x0 - 1, 0);
    else if (lane == 1 || lane == 2) ho = (uint32_t)jc * syu + (uint32_t)irc;
    else if (lane == 3) ho = xpml_r ? xoff(sxr, jc) : 0u;
    else ho = o;
    r.halo = ld<T>(hbase, ho);
    r.ez = ld<T>(a.ez_a, o);
    r.hx = ld<T>(a.hx_a, o);
    r.hy = ld<T>(a.hy_a, o);
    const int syj = g.has_pml ? slab(jc, ny, py0, py1) : -1;
    r.psx = ld<T>(a.psh_y_a, syj >= 0 ? yoff(ic, syj) : 0u);   // Hx: y slabs
    r.psy = ld<T>(a.psh_x_a, xpml ? xoff(sxs, jc) : 0u);       // Hy: x slabs
  };
  auto load_e = [&](int jj, TmRowE<T, CK>& r) __attribute__((always_inline)) {
    const int jc = min(jj, ny - 1);
    const uint32_t o = (uint32_t)jc * syu + (uint32_t)ic;
    r.cb = ld<CbT>(a.cb, o);
    const int syj = g.has_pml ? slab(jc, ny, py0, py1) : -1;
    r.pex = ld<T>(a.pse_x, xpml ? xoff(sxs, jc) : 0u);
    r.pey = ld<T>(a.pse_y, syj >= 0 ? yoff(ic, syj) : 0u);
  };

  const int r0 = tile * a.chunk, r1 = min(r0 + a.chunk, ny);
  const int jend = min(r1, ny - 1);
  const int elast = max(jend - 1, r0);
  T ezp = (r0 >= 1) ? ld<T>(a.ez_a, (uint32_t)(r0 - 1) * syu + (uint32_t)ic) : (T)0;
  T hxp = 0, hyp = 0, hyp_r = 0;

  auto body = [&](int j, const TmRowH<T>& cur, TmRowH<T>& nxt, const TmRowE<T, CK>& ecur,
                  TmRowE<T, CK>& enxt) __attribute__((always_inline)) {
    load_h(min(j + 2, jend), nxt);
    load_e(min(j + 1, elast), enxt);
    const uint32_t o = (uint32_t)j * syu + (uint32_t)i;
    const bool own = inx && j < r1;
    const int syj = g.has_pml ? slab(j, ny, py0, py1) : -1;
    const T ezc = cur.ez;

    // Hx (i < nx-1): curl = dEz/dy - 0, backward difference
    T d1 = (ezc - ezp) * idy;
    if (syj >= 0) {
      const T ps = ybh[j] * cur.psx + ych[j] * d1;
      if (own && i < nx - 1) st<T>(a.psh_y_b, yoff(i, syj), ps);
      d1 = d1 * yikh[j] + ps;
    }
    const T hxn = (inx && i < nx - 1) ? cur.hx - db * (d1 - (T)0) : (T)0;

    // Hy (j < ny-1): curl = 0 - dEz/dx, Ez(i-1) from the lane below (lane
    // 0: the halo value, zero left of the grid)
    const T ezl0 = x0 > 0 ? lane_val(cur.halo, 0) : (T)0;
    const T ezs = __shfl_up(ezc, 1);
    const T ezl = lane == 0 ? ezl0 : (i > 0 ? ezs : (T)0);
    T d2 = (ezc - ezl) * idx;
    if (xpml) {
      const T ps = bhx * cur.psy + chx * d2;
      if (own && j < ny - 1) st<T>(a.psh_x_b, xoff(sxs, j), ps);
      d2 = d2 * ikhx + ps;
    }
    const T hyn = (inx && j < ny - 1) ? cur.hy - db * ((T)0 - d2) : (T)0;
    // Hy(n+1/2) at x0+64 for lane 63's E update (not stored)
    T d2r = (lane_val(cur.halo, 1) - lane_val(ezc, 63)) * idx;
    if (xpml_r) {
      const T ps = bhx_r * lane_val(cur.halo, 3) + chx_r * d2r;
      d2r = d2r * ikhx_r + ps;
    }
    const T hyn_r = (rin && j < ny - 1) ? lane_val(cur.halo, 2) - db * ((T)0 - d2r) : (T)0;
    if (own) {
      st<T>(a.hx_b, o, hxn);
      st<T>(a.hy_b, o, hyn);
    }

    if (j > r0) {
      // Ez(i, jj) (i < nx-1, jj < ny-1): curl = dHy/dx - dHx/dy with the
      // new H of rows jj and jj+1
      const int jj = j - 1;
      const uint32_t oe = o - syu;
      const int sye = g.has_pml ? slab(jj, ny, py0, py1) : -1;
      const T hys1 = __shfl_down(hyp, 1);
      const T hyr = lane == 63 ? hyp_r : hys1;
      const bool valid = inx && i < nx - 1 && jj < ny - 1;
      T e1 = (hyr - hyp) * idx;
      if (xpml) {
        const T ps = bex * ecur.pex + cex * e1;
        if (valid) st<T>(a.pse_x, xoff(sxs, jj), ps);
        e1 = e1 * ikex + ps;
      }
      T e2 = (hxn - hxp) * idy;
      if (sye >= 0) {
        const T ps = ybe[jj] * ecur.pey + yce[jj] * e2;
        if (valid) st<T>(a.pse_y, yoff(i, sye), ps);
        e2 = e2 * yike[jj] + ps;
      }
      T cb;
      if constexpr (CK == PM_COEFF_CB_F16) {
        __half_raw hr;
        hr.x = ecur.cb;
        cb = cbs * (T)__half2float(__half(hr));
      } else {
        cb = ecur.cb;
      }
      const T ezn = valid ? ezp + cb * (e1 - e2) : (T)0;
      if (inx) st<T>(a.ez_b, oe, ezn);
    }
    ezp = ezc;
    hxp = hxn;
    hyp = hyn;
    hyp_r = hyn_r;
  };

  TmRowH<T> h0, h1, h2;
  TmRowE<T, CK> e0, e1, e2;
  load_h(r0, h0);
  load_h(min(r0 + 1, jend), h1);
  e0 = TmRowE<T, CK>{};
  load_e(r0, e1);
  for (int j = r0;;) {
    body(j, h0, h2, e0, e2);
    if (++j > jend) break;
    body(j, h1, h0, e1, e0);
    if (++j > jend) break;
    body(j, h2, h1, e2, e1);
    if (++j > jend) break;
  }
  // the last row of the grid holds no Ez sample: its B copy stays zero (it
  // is zero in A as well), written by the tile that owns it
  if (r1 == ny && ny >= 1 && inx) st<T>(a.ez_b, (uint32_t)(ny - 1) * syu + (uint32_t)i, (T)0);
}

}  // namespace

int launch_tm2d_fused(const KGrid& g, int precision, int ck, const Tm2dBufs& b, int chunk,
                      hipStream_t s) {
  if (g.ndim != 2 || chunk < 1 || g.n_ade) return -1;
  if (ck != PM_COEFF_CB && !(ck == PM_COEFF_CB_F16 && precision == PM_FP32)) return -1;
  if ((uint64_t)g.sy * (uint64_t)g.ny >= (1ull << 31)) return -1;
  Tm2dArgs a{};
  a.ez_a = b.Ei;
  a.hx_a = b.Hi[0];
  a.hy_a = b.Hi[1];
  a.ez_b = b.Eo;
  a.hx_b = b.Ho[0];
  a.hy_b = b.Ho[1];
  // axes without CPML read element 0 of a field array (never used) and
  // store nothing
  a.psh_x_a = b.psh_in[0] ? b.psh_in[0] : b.Ei;
  a.psh_y_a = b.psh_in[1] ? b.psh_in[1] : b.Ei;
  a.psh_x_b = b.psh_out[0] ? b.psh_out[0] : b.Eo;
  a.psh_y_b = b.psh_out[1] ? b.psh_out[1] : b.Eo;
  a.pse_x = g.psi_e[0][1] ? g.psi_e[0][1] : b.Eo;
  a.pse_y = g.psi_e[1][0] ? g.psi_e[1][0] : b.Eo;
  a.cb = g.cb[2];
  a.chunk = chunk;
  const int nstrip = (g.nx + 63) / 64;
  const int ntile = (g.ny + chunk - 1) / chunk;
  const int n_items = nstrip * ntile;
  const int nwg = (n_items + kWaves2d - 1) / kWaves2d;
  const dim3 grid(nwg), block(64 * kWaves2d);
  if (precision == PM_FP64)
    hipLaunchKernelGGL((k_tm2d_fused<double, PM_COEFF_CB>), grid, block, 0, s, g, a, n_items);
  else if (ck == PM_COEFF_CB)
    hipLaunchKernelGGL((k_tm2d_fused<float, PM_COEFF_CB>), grid, block, 0, s, g, a, n_items);
  else
    hipLaunchKernelGGL((k_tm2d_fused<float, PM_COEFF_CB_F16>), grid, block, 0, s, g, a, n_items);
  return 0;
}

}  // namespace pm
