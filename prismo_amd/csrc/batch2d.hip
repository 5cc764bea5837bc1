// Batched 2-D TM stepping of many uneven grids (include/pmbatch.h).
//
// One launch advances every design by one step.  The work list holds one
// tile per wave: (design, x0, r0, r1) = a strip of 64 x cells [x0, x0+64),
// x0 a multiple of 64 (rows are padded to 64 cells, so every row access of
// a wave is two whole 128-byte lines), and the rows [r0, r1).  The wave
// marches up its rows once, carrying row j-1 in registers:
//   H(n+1/2) at row j   from Ez(j), Ez(j-1) (registers) and Ez(i-1) (shuffle);
//   E(n+1)   at row j-1 from Hx(j), Hx(j-1), Hy(i), Hy(i+1) (registers/shuffle).
// The two values from outside the strip come from one extra load per row
// (lanes 0..3: Ez(x0-1), and Ez, Hy and the Hy CPML state at x0+64): lane 0
// takes Ez(x0-1) as its left neighbour, and Hy(n+1/2) at x0+64, which lane
// 63's E update needs, is recomputed from them with the arithmetic of the
// next strip's lane 0.  So every array is read once and written once per
// step (plus one halo row per tile).  Reads come from field set
// `par`, writes go to set par^1, so neighbouring tiles never see a
// half-updated value; the H-side CPML memory is ping-ponged for the same
// reason, the E-side memory is touched only by the owner and stays in place.
// Sources are added after the E update, samples are taken after the sources,
// as src/prismo/core/simulation.py:147-164 orders them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/pmbatch.h"
#include "pm_internal.h"

namespace pm {
namespace {

constexpr int kStrip = 64;    // cells per wave
constexpr int kMaxWaves = 8;  // waves per workgroup: 1, 2, 4 (default) or 8

struct KWork {
  int d, x0, r0, r1;
};

struct KBatch {
  const pm_batch_design* des;
  const KWork* work;
  int n_work;
  int ncols, sample_T;
  int64_t* clock;
  int32_t* status;  // per design: bit 0 set once an Ez value was not finite
  int waves;        // waves per workgroup (the launch's blockDim.x / 64)
  int xcd;          // 1: XCD-grouped workgroup order, 0: launch order
  double idx, idy;
  const double* tab_e;
  const double* tab_h;
  int64_t table_step0;
  void* X;
};

// Loads of one row for the H update (fields, H-side CPML memory and the
// y profile at nodes) and of one row for the E update (Cb, E-side CPML
// memory, y profile at half nodes).  They are issued one iteration before
// their use; the loop below alternates two register sets instead of copying
// so no wait lands on the loop's back edge.
template <typename T>
struct RowH {
  T ez, hx, hy, psx, psy;
  T halo;  // lanes 0..3: Ez(x0-1), Ez(x0+64), Hy(x0+64), Hy CPML state at x0+64
};

// Element e of an array in global memory: a uniform base plus a 32-bit
// byte offset, so one offset register serves every array of a row (the
// host checks that a design's arrays fit 32-bit byte offsets).
template <typename T>
__device__ __forceinline__ T ldg(const void* base, uint32_t e) {
  return *(gptr<const T>)((gptr<const char>)base + e * (uint32_t)sizeof(T));
}

// Stores are non-temporal: the other field set is read one step later,
// after every design's step, so caching it only displaces reused lines.
template <typename T>
__device__ __forceinline__ void stg(void* base, uint32_t e, T v) {
  __builtin_nontemporal_store(v, (gptr<T>)((gptr<char>)base + e * (uint32_t)sizeof(T)));
}

// Value of v in lane l (l uniform), for the whole wave.
template <typename T>
__device__ __forceinline__ T lane_value(T v, int l) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
  } else {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
  }
}

// Bit 31 set exactly for Inf and NaN: the exponent field plus one carries
// into the sign position only when the field is all ones.  Integer VALU
// work, OR-ed into one register per lane (a compare would hold a lane mask
// in scalar registers across the row loop), and no floating-point mode can
// fold it away.
template <typename T>
__device__ __forceinline__ uint32_t nonfinite_bit(T v) {
  if constexpr (sizeof(T) == 4)
    return (__builtin_bit_cast(uint32_t, v) & 0x7f800000u) + 0x00800000u;
  else
    return ((uint32_t)(__builtin_bit_cast(uint64_t, v) >> 32) & 0x7ff00000u) + 0x00100000u;
}

template <typename T, int CK>
struct RowE {
  using CbT = typename std::conditional<CK == PM_COEFF_CB_F16, unsigned short, T>::type;
  CbT cb;
  T pex, pey;
};

template <typename T, int CK>
__global__ __launch_bounds__(64 * kMaxWaves) void k_batch_tm(const KBatch k, int64_t step_arg,
                                                             int use_clock, int s_rel,
                                                             int par) {
  const int lane = threadIdx.x & 63;
  // the wave index is uniform: make that visible so the tile and design
  // descriptors are scalar loads
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  // workgroups go round-robin over the 8 XCDs (b % 8); in each group of 64,
  // XCD x takes work groups 8x .. 8x+7, so neighbouring strips (which read
  // the same cache lines at their edges) share one L2
  const int b = blockIdx.x, nb = gridDim.x;
  const int wg = k.xcd && b < (nb & ~63) ? (b & ~63) + (b & 7) * 8 + ((b >> 3) & 7) : b;
  const int item = wg * k.waves + wv;
  if (item >= k.n_work) return;  // whole wave
  const KWork w = k.work[item];
  const pm_batch_design& D = k.des[w.d];
  const int64_t step = use_clock ? (*k.clock + s_rel) : step_arg;
  const int nx = D.nx, ny = D.ny;
  const int64_t sy = D.sy;
  const int x0 = w.x0;
  const int i = x0 + lane;
  const bool inx = i < nx;
  const bool own = inx;
  const int q = par, qn = par ^ 1;
  // the design table holds generic pointers; they all address device
  // memory, and saying so gives global (not flat) loads, whose counters the
  // compiler can wait on per load
  const void* const EzA = D.ez[q];
  const void* const HxA = D.hx[q];
  const void* const HyA = D.hy[q];
  void* const EzB = D.ez[qn];
  void* const HxB = D.hx[qn];
  void* const HyB = D.hy[qn];
  const void* const psHxA = D.psi_hx[q];
  const void* const psHyA = D.psi_hy[q];
  void* const psHxB = D.psi_hx[qn];
  void* const psHyB = D.psi_hy[qn];
  void* const psEx = D.psi_ex;
  void* const psEy = D.psi_ey;
  const gptr<const T> prof = (gptr<const T>)D.prof;
  // y profiles, 8 per row: read-only for the launch, so scalar loads when
  // a row needs them (no vector registers held for them across the
  // prefetch)
  using CT = const T __attribute__((address_space(4)))*;
  const CT pyr = (CT)(prof + 6 * (int64_t)nx);

  // x profiles of this lane (identity outside the slabs)
  const bool xpml = inx && (i < D.pml[0] || i >= nx - D.pml[1]);
  T bex = 1, cex = 0, ikex = 1, bhx = 1, chx = 0, ikhx = 1;
  if (xpml) {
    bex = prof[i];
    cex = prof[nx + i];
    ikex = prof[2 * nx + i];
    bhx = prof[3 * nx + i];
    chx = prof[4 * nx + i];
    ikhx = prof[5 * nx + i];
  }
  // the same for the cell right of the strip, x0+64 (uniform)
  const int ir = x0 + kStrip;
  const bool rin = ir < nx;
  const int irc = min(ir, nx - 1);
  const bool xpml_r = rin && (ir < D.pml[0] || ir >= nx - D.pml[1]);
  T bhx_r = 1, chx_r = 0, ikhx_r = 1;
  if (xpml_r) {
    bhx_r = prof[3 * nx + ir];
    chx_r = prof[4 * nx + ir];
    ikhx_r = prof[5 * nx + ir];
  }
  const T db = (T)D.db;
  const T idx = (T)k.idx, idy = (T)k.idy;
  const int64_t trow = step - k.table_step0;
  const T se = (T)k.tab_e[trow];
  const T sh = (T)k.tab_h[trow];
  const gptr<T> Xrow = (gptr<T>)k.X + (step % k.sample_T) * (int64_t)k.ncols;
  // everything the row loop needs from the design table, read once: the
  // loop's stores could alias the table as far as the compiler knows
  const int pml_y0 = D.pml[2], pml_y1 = ny - D.pml[3];
  const bool srcE = i == D.src_e[0];
  const bool srcH = i == D.src_h[0];
  const int se0 = D.src_e[1], se1 = D.src_e[2];
  const int sh0 = D.src_h[1], sh1 = D.src_h[2];
  const T cbs = (T)D.cb_scale;
  using CbT = typename RowE<T, CK>::CbT;
  const void* const cbp = D.cb;
  // sample lines at this lane's column (bit l: line l)
  unsigned lmask = 0;
  const int nl = D.n_lines;
  for (int l = 0; l < nl; ++l)
    if (D.line[l].i == i) lmask |= 1u << l;
  if (!own) lmask = 0;
  const bool cbx = inx && i < nx - 1;

  // Every load of the row loop is issued on every path, from a clamped
  // in-range position whose value is masked where it is used (lanes outside
  // the design, rows past the last, cells outside a CPML slab), so the
  // compiler's memory-counter waits do not cover the prefetched rows.
  const int ic = min(i, nx - 1);
  // the halo load: lane 1 Ez, lane 2 Hy, lane 3 the Hy CPML state (element
  // 0 outside the x slabs) at x0+64; every other lane Ez(x0-1) (element 0
  // of the row for the first strip)
  const gptr<const T> hbase = (gptr<const T>)(lane == 2 ? HyA : (lane == 3 ? psHyA : EzA));
  const int hcol = (lane == 1 || lane == 2 || (lane == 3 && xpml_r)) ? irc : max(x0 - 1, 0);
  const bool hrow = !(lane == 3 && !xpml_r);  // false: element 0
  const uint32_t syu = (uint32_t)sy;
  auto load_h = [&](int j, RowH<T>& r) __attribute__((always_inline)) {
    const int jc = min(j, ny - 1);
    const uint32_t o = (uint32_t)jc * syu + (uint32_t)ic;
    r.halo = hbase[hrow ? (int64_t)jc * sy + hcol : 0];
    r.ez = ldg<T>(EzA, o);
    r.hx = ldg<T>(HxA, o);
    r.hy = ldg<T>(HyA, o);
    // the CPML states are used in the slabs only: elsewhere every lane
    // reads element 0 (one cache line per wave, no extra HBM traffic) and
    // the value is not used
    r.psx = ldg<T>(psHxA, (j < pml_y0 || j >= pml_y1) ? o : 0u);
    r.psy = ldg<T>(psHyA, xpml ? o : 0u);
  };
  auto load_e = [&](int jj, RowE<T, CK>& r) __attribute__((always_inline)) {
    const int jc = min(jj, ny - 1);
    const uint32_t o = (uint32_t)jc * syu + (uint32_t)ic;
    r.cb = ldg<CbT>(cbp, o);
    r.pex = ldg<T>(psEx, xpml ? o : 0u);
    r.pey = ldg<T>(psEy, (jj < pml_y0 || jj >= pml_y1) ? o : 0u);
  };

  const int r0 = w.r0, r1 = w.r1;
  const int jend = min(r1, ny - 1);
  const int elast = max(jend - 1, r0);  // last E row the tile updates (or r0)
  T ezp = (r0 >= 1 && inx) ? ldg<T>(EzA, (uint32_t)(r0 - 1) * syu + (uint32_t)i) : (T)0;
  T hxp = 0, hyp = 0, hyp_r = 0;
  uint32_t bad = 0;  // bit 31: this lane computed an Inf or NaN Ez

  // one row: H(n+1/2) at row j from `cur`, E(n+1) at row j-1 from `ecur`;
  // first loads row j+2 into `nxt` and the E row j+1 into `enxt` (two rows
  // ahead: the loads of two rows are in flight while one is computed)
  auto body = [&](int j, const RowH<T>& cur, RowH<T>& nxt, const RowE<T, CK>& ecur,
                  RowE<T, CK>& enxt) __attribute__((always_inline)) {
    // rows past the tile are not used: the clamp makes those loads re-read
    // a row already in the cache instead of fetching the next tile's rows
    load_h(min(j + 2, jend), nxt);
    load_e(min(j + 1, elast), enxt);
    const uint32_t o = (uint32_t)j * syu + (uint32_t)i;
    const bool store_h = own && j < r1;
    const bool yp = j < pml_y0 || j >= pml_y1;
    const T ezc = inx ? cur.ez : (T)0;

    // Hx(i, j) -= Db * dEz/dy (backward difference, Ez(j = -1) = 0)
    T dy = (ezc - ezp) * idy;
    if (yp) {
      const CT q = pyr + 8 * j;
      const T ps = q[3] * cur.psx + q[4] * dy;
      if (store_h) stg<T>(psHxB, o, ps);
      dy = dy * q[5] + ps;
    }
    const T hxn = (inx && i < nx - 1) ? cur.hx - db * dy : (T)0;

    // Hy(i, j) += Db * dEz/dx, Ez(i-1) from the lane below (lane 0: the
    // halo value, zero left of the grid)
    const T ezh = x0 > 0 ? lane_value(cur.halo, 0) : (T)0;
    const T ezs = __shfl_up(ezc, 1);
    const T ezl = lane == 0 ? ezh : ezs;
    T dx = (ezc - ezl) * idx;
    if (xpml) {
      const T ps = bhx * cur.psy + chx * dx;
      if (store_h) stg<T>(psHyB, o, ps);
      dx = dx * ikhx + ps;
    }
    const T hyn = (inx && j < ny - 1) ? cur.hy + db * dx : (T)0;
    // Hy(n+1/2) at x0+64 for lane 63's E update: the next strip's lane 0
    // arithmetic on the halo values (not stored)
    T dxr = (lane_value(cur.halo, 1) - lane_value(ezc, 63)) * idx;
    if (xpml_r) {
      const T ps = bhx_r * lane_value(cur.halo, 3) + chx_r * dxr;
      dxr = dxr * ikhx_r + ps;
    }
    const T hyn_r = (rin && j < ny - 1) ? lane_value(cur.halo, 2) + db * dxr : (T)0;

    if (store_h) {
      stg<T>(HxB, o, hxn);
      const T hys = (srcH && j >= sh0 && j < sh1) ? hyn + sh : hyn;
      stg<T>(HyB, o, hys);
      if (lmask) {
        // a constant trip count (lmask holds no bit at or past nl): stores
        // in a loop of run-time length would make the compiler wait for
        // every access in flight, the row prefetch included
#pragma unroll
        for (int l = 0; l < PM_BATCH_MAX_LINES; ++l) {
          if (!((lmask >> l) & 1u)) continue;
          const pm_batch_line L = D.line[l];
          if (L.comp != 0 && j >= L.j0 && j < L.j1)
            Xrow[L.col + (j - L.j0)] = L.comp == 2 ? hys : hxn;
        }
      }
    }

    if (j > r0) {
      // Ez(i, jj) += Cb (dHy/dx - dHx/dy) with the new H of rows jj, jj+1
      const int jj = j - 1;
      const uint32_t oe = o - syu;
      const T hys1 = __shfl_down(hyp, 1);
      const T hyr = lane == 63 ? hyp_r : hys1;
      const bool valid = cbx && jj < ny - 1;
      T tx = (hyr - hyp) * idx;
      if (xpml && valid) {
        const T ps = bex * ecur.pex + cex * tx;
        if (own) stg<T>(psEx, oe, ps);
        tx = tx * ikex + ps;
      }
      T ty = (hxn - hxp) * idy;
      if ((jj < pml_y0 || jj >= pml_y1) && valid) {
        const CT q = pyr + 8 * jj;
        const T ps = q[0] * ecur.pey + q[1] * ty;
        if (own) stg<T>(psEy, oe, ps);
        ty = ty * q[2] + ps;
      }
      T cb;
      if constexpr (CK == PM_COEFF_CB_F16) {
        __half_raw hr;
        hr.x = ecur.cb;
        cb = cbs * (T)__half2float(__half(hr));
      } else {
        cb = ecur.cb;
      }
      T ezn = (T)0;
      if (valid) {
        ezn = ezp + cb * (tx - ty);
        if (srcE && jj >= se0 && jj < se1) ezn += se;
        bad |= nonfinite_bit(ezn);
      }
      if (own) {
        stg<T>(EzB, oe, ezn);
        if (lmask) {
#pragma unroll
          for (int l = 0; l < PM_BATCH_MAX_LINES; ++l) {
            if (!((lmask >> l) & 1u)) continue;
            const pm_batch_line L = D.line[l];
            if (L.comp == 0 && jj >= L.j0 && jj < L.j1) Xrow[L.col + (jj - L.j0)] = ezn;
          }
        }
      }
    }
    ezp = ezc;
    hxp = hxn;
    hyp = hyn;
    hyp_r = hyn_r;
  };

  // three register sets in rotation: iteration j computes from set
  // (j - r0) % 3 and loads into set (j - r0 + 2) % 3
  RowH<T> h0, h1, h2;
  RowE<T, CK> e0, e1, e2;
  load_h(r0, h0);
  load_h(min(r0 + 1, jend), h1);
  e0 = RowE<T, CK>{};  // E row r0 - 1: not updated by this tile
  load_e(r0, e1);
  for (int j = r0;;) {
    body(j, h0, h2, e0, e2);
    if (++j > jend) break;
    body(j, h1, h0, e1, e0);
    if (++j > jend) break;
    body(j, h2, h1, e2, e1);
    if (++j > jend) break;
  }
  // a diverged design is flagged once per tile (rare: a vector atomic from
  // the lanes that saw it)
  if (bad >> 31) atomicOr(k.status + w.d, 1);
}

__global__ void k_clock_set(int64_t* clock, int64_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *clock = v;
}

__global__ void k_clock_add(int64_t* clock, int64_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *clock += v;
}

}  // namespace
}  // namespace pm

using namespace pm;

struct pm_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  bool configured = false;
  pm_batch_plan plan{};
  std::vector<pm_batch_design> des;
  std::vector<KWork> work;
  KBatch kb{};
  KDft dft{};
  int64_t step = 0;   // absolute index of the next step
  int parity = 0;
  int64_t pend0 = 0;  // first step of the pending sample block
  int pend_n = 0;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  int waves = 4;  // launch shape of the step kernel (pm_batch_set_tuning)
  int xcd = 1;
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> spans[3];
  std::string err;
};

namespace {

int bfail(pm_batch* b, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (b) b->err = buf;
  return code;
}

int bhip(pm_batch* b, hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  return bfail(b, (int)e, "%s: %s", what, hipGetErrorString(e));
}

struct BSpan {
  pm_batch* b;
  int which;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  BSpan(pm_batch* bb, int w) : b(bb), which(w) {
    if (b->timing && hipEventCreateWithFlags(&e0, hipEventDisableSystemFence) == hipSuccess &&
        hipEventCreateWithFlags(&e1, hipEventDisableSystemFence) == hipSuccess)
      (void)hipEventRecord(e0, b->stream);
  }
  ~BSpan() {
    if (e0 && e1) {
      (void)hipEventRecord(e1, b->stream);
      b->spans[which].push_back({e0, e1});
    }
  }
};

int64_t align256(int64_t v) { return (v + 255) & ~(int64_t)255; }

// Workspace: the step clock at offset 0, then the design table, the tile
// list and one status word per design.
void layout(const pm_batch_plan& p, int64_t* des_off, int64_t* work_off, int64_t* status_off,
            int64_t* total, int64_t n_work) {
  *des_off = 256;
  *work_off = align256(*des_off + (int64_t)sizeof(pm_batch_design) * p.n_design);
  *status_off = align256(*work_off + (int64_t)sizeof(KWork) * n_work);
  *total = align256(*status_off + (int64_t)sizeof(int32_t) * p.n_design);
}

void build_work(const pm_batch_plan& p, std::vector<KWork>& work) {
  work.clear();
  const int ch = p.chunk;
  for (int d = 0; d < p.n_design; ++d) {
    const pm_batch_design& D = p.design[d];
    for (int r0 = 0; r0 < D.ny; r0 += ch)
      for (int x0 = 0; x0 < D.nx; x0 += kStrip)
        work.push_back(KWork{d, x0, r0, std::min(r0 + ch, D.ny)});
  }
}

int validate(pm_batch* b, const pm_batch_plan& p) {
  if (p.abi_version != PM_ABI_VERSION) return bfail(b, PM_EINVAL, "ABI version mismatch");
  if (p.precision != PM_FP64 && p.precision != PM_FP32)
    return bfail(b, PM_EINVAL, "precision must be PM_FP64 or PM_FP32");
  if (p.coeff_kind != PM_COEFF_CB && p.coeff_kind != PM_COEFF_CB_F16)
    return bfail(b, PM_EINVAL, "coeff_kind must be CB or CB_F16");
  if (p.n_design < 1 || !p.design) return bfail(b, PM_EINVAL, "no designs");
  if (!(p.dx > 0) || !(p.dy > 0) || !(p.dt > 0)) return bfail(b, PM_EINVAL, "bad dx/dy/dt");
  if (p.chunk < 2) return bfail(b, PM_EINVAL, "chunk must be >= 2");
  if (p.sample_T < 1 || p.ncols < 0 || !p.X) return bfail(b, PM_EINVAL, "bad sample ring");
  if (!p.src_e_table || !p.src_h_table || p.table_len < 1)
    return bfail(b, PM_EINVAL, "source tables required");
  if (p.nfreq < 0) return bfail(b, PM_EINVAL, "bad nfreq");
  if (p.nfreq > 0) {
    if (!p.W_re || !p.W_im || !p.dft.acc_re || !p.dft.acc_im)
      return bfail(b, PM_EINVAL, "spectra need W tables and accumulators");
    if (p.dft.col0 < 0 || p.dft.col1 > p.ncols || p.dft.col0 >= p.dft.col1 ||
        p.dft.f0 < 0 || p.dft.f1 > p.nfreq || p.dft.f0 >= p.dft.f1)
      return bfail(b, PM_EINVAL, "DFT group outside the sample columns / frequencies");
    if (p.precision == PM_FP32 && (!p.dft.cmp_re || !p.dft.cmp_im))
      return bfail(b, PM_EINVAL, "FP32 spectra need compensation arrays");
  }
  for (int d = 0; d < p.n_design; ++d) {
    const pm_batch_design& D = p.design[d];
    if (D.nx < 2 || D.ny < 2 || D.sy < D.nx) return bfail(b, PM_EINVAL, "design %d: bad size", d);
    // the kernel addresses a design's arrays with 32-bit byte offsets
    if ((int64_t)D.ny * D.sy * (p.precision == PM_FP64 ? 8 : 4) >= (int64_t)1 << 32)
      return bfail(b, PM_EINVAL, "design %d: %d x %lld cells exceed 32-bit offsets", d, D.ny,
                   (long long)D.sy);
    for (int s = 0; s < 2; ++s)
      if (!D.ez[s] || !D.hx[s] || !D.hy[s] || !D.psi_hx[s] || !D.psi_hy[s])
        return bfail(b, PM_EINVAL, "design %d: missing field or CPML arrays", d);
    if (!D.psi_ex || !D.psi_ey || !D.cb || !D.prof)
      return bfail(b, PM_EINVAL, "design %d: missing arrays", d);
    if (D.pml[0] < 0 || D.pml[1] < 0 || D.pml[2] < 0 || D.pml[3] < 0 ||
        D.pml[0] + D.pml[1] > D.nx || D.pml[2] + D.pml[3] > D.ny)
      return bfail(b, PM_EINVAL, "design %d: bad CPML layers", d);
    if (D.n_lines < 0 || D.n_lines > PM_BATCH_MAX_LINES)
      return bfail(b, PM_EINVAL, "design %d: too many sample lines", d);
    for (int l = 0; l < D.n_lines; ++l) {
      const pm_batch_line& L = D.line[l];
      if (L.comp < 0 || L.comp > 2 || L.i < 0 || L.i >= D.nx || L.j0 < 0 || L.j1 > D.ny ||
          L.j0 > L.j1 || L.col < 0 || L.col + (L.j1 - L.j0) > p.ncols)
        return bfail(b, PM_EINVAL, "design %d: bad sample line %d", d, l);
    }
    // Ez exists on (nx-1) x (ny-1), Hy on nx x (ny-1)
    if (D.src_e[1] < 0 || D.src_e[1] > D.src_e[2] ||
        (D.src_e[2] > D.src_e[1] && (D.src_e[0] < 0 || D.src_e[0] >= D.nx - 1 ||
                                     D.src_e[2] > D.ny - 1)))
      return bfail(b, PM_EINVAL, "design %d: Ez source line outside Ez", d);
    if (D.src_h[1] < 0 || D.src_h[1] > D.src_h[2] ||
        (D.src_h[2] > D.src_h[1] && (D.src_h[0] < 0 || D.src_h[0] >= D.nx ||
                                     D.src_h[2] > D.ny - 1)))
      return bfail(b, PM_EINVAL, "design %d: Hy source line outside Hy", d);
  }
  return 0;
}

template <typename T>
int launch_step(pm_batch* b, int64_t step_arg, int use_clock, int s_rel, int par) {
  const int wv = b->kb.waves;
  const int nb = ((int)b->work.size() + wv - 1) / wv;
  if (b->plan.coeff_kind == PM_COEFF_CB_F16)
    hipLaunchKernelGGL((k_batch_tm<T, PM_COEFF_CB_F16>), dim3(nb), dim3(64 * wv), 0,
                       b->stream, b->kb, step_arg, use_clock, s_rel, par);
  else
    hipLaunchKernelGGL((k_batch_tm<T, PM_COEFF_CB>), dim3(nb), dim3(64 * wv), 0,
                       b->stream, b->kb, step_arg, use_clock, s_rel, par);
  return bhip(b, hipGetLastError(), "batch step launch");
}

int step_once(pm_batch* b, int64_t step, int use_clock, int s_rel, int par) {
  return b->plan.precision == PM_FP64 ? launch_step<double>(b, step, use_clock, s_rel, par)
                                      : launch_step<float>(b, step, use_clock, s_rel, par);
}

int flush_block(pm_batch* b, int64_t first, int rows, const int64_t* wclock) {
  const pm_batch_plan& p = b->plan;
  if (p.nfreq == 0 || rows == 0) return 0;
  const int esz = p.precision == PM_FP64 ? 8 : 4;
  const char* X0 = (const char*)p.X + (first % p.sample_T) * (int64_t)p.ncols * esz;
  const int64_t wcol = wclock ? p.table_step0 : first - p.table_step0;
  launch_dft(b->dft, X0, p.ncols, p.W_re, p.W_im, p.table_len, wcol, rows, p.precision,
             b->stream, wclock);
  return bhip(b, hipGetLastError(), "batch DFT launch");
}

int capture(pm_batch* b, int par) {
  const pm_batch_plan& p = b->plan;
  hipGraph_t g = nullptr;
  int r = bhip(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal),
               "begin capture");
  if (r) return r;
  for (int s = 0; s < p.sample_T && !r; ++s) r = step_once(b, 0, 1, s, par ^ (s & 1));
  if (!r) r = flush_block(b, 0, p.sample_T, b->kb.clock);
  if (!r) {
    hipLaunchKernelGGL(k_clock_add, dim3(1), dim3(64), 0, b->stream, b->kb.clock,
                       (int64_t)p.sample_T);
    r = bhip(b, hipGetLastError(), "clock launch");
  }
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (r) {
    if (g) (void)hipGraphDestroy(g);
    return r;
  }
  r = bhip(b, e, "end capture");
  if (r) return r;
  r = bhip(b, hipGraphInstantiate(&b->graph[par], g, nullptr, nullptr, 0), "instantiate");
  (void)hipGraphDestroy(g);
  return r;
}

void drop_graphs(pm_batch* b) {
  for (auto& g : b->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

}  // namespace

extern "C" {

int pm_batch_create(int device, void* stream, pm_batch** out) {
  if (!out) return PM_EINVAL;
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) return (int)e;
  if (device < 0 || device >= n) return PM_EINVAL;
  e = hipSetDevice(device);
  if (e != hipSuccess) return (int)e;
  pm_batch* b = new pm_batch();
  b->device = device;
  b->stream = (hipStream_t)stream;
  *out = b;
  return 0;
}

int pm_batch_destroy(pm_batch* b) {
  if (!b) return PM_EINVAL;
  (void)hipSetDevice(b->device);
  drop_graphs(b);
  for (auto& v : b->spans)
    for (auto& ab : v) {
      (void)hipEventDestroy(ab.first);
      (void)hipEventDestroy(ab.second);
    }
  delete b;
  return 0;
}

const char* pm_batch_last_error(pm_batch* b) { return b ? b->err.c_str() : "null handle"; }

int pm_batch_workspace_bytes(const pm_batch_plan* p, int64_t* bytes) {
  if (!p || !bytes || p->n_design < 1 || !p->design || p->chunk < 2) return PM_EINVAL;
  std::vector<KWork> w;
  build_work(*p, w);
  int64_t a, c, st, t;
  layout(*p, &a, &c, &st, &t, (int64_t)w.size());
  *bytes = t;
  return 0;
}

int pm_batch_configure(pm_batch* b, const pm_batch_plan* p) {
  if (!b || !p) return PM_EINVAL;
  (void)hipSetDevice(b->device);
  b->configured = false;
  drop_graphs(b);
  int r = validate(b, *p);
  if (r) return r;
  std::vector<KWork> work;
  build_work(*p, work);
  int64_t des_off, work_off, status_off, total;
  layout(*p, &des_off, &work_off, &status_off, &total, (int64_t)work.size());
  if (!p->work || p->work_bytes < total)
    return bfail(b, PM_EINVAL, "workspace of %lld bytes required", (long long)total);
  b->plan = *p;
  b->des.assign(p->design, p->design + p->n_design);
  b->plan.design = b->des.data();
  b->work = std::move(work);
  char* ws = (char*)p->work;
  KBatch& k = b->kb;
  k.clock = (int64_t*)ws;
  k.des = (const pm_batch_design*)(ws + des_off);
  k.work = (const KWork*)(ws + work_off);
  k.status = (int32_t*)(ws + status_off);
  k.n_work = (int)b->work.size();
  k.waves = b->waves;
  k.xcd = b->xcd;
  k.ncols = p->ncols;
  k.sample_T = p->sample_T;
  k.idx = 1.0 / p->dx;
  k.idy = 1.0 / p->dy;
  k.tab_e = p->src_e_table;
  k.tab_h = p->src_h_table;
  k.table_step0 = p->table_step0;
  k.X = p->X;
  b->dft = KDft{p->dft.col0, p->dft.col1, p->dft.f0, p->dft.f1, p->dft.acc_re,
                p->dft.acc_im, p->dft.cmp_re, p->dft.cmp_im};
  r = bhip(b, hipMemcpyAsync((void*)k.des, b->des.data(),
                             sizeof(pm_batch_design) * b->des.size(), hipMemcpyHostToDevice,
                             b->stream), "design table upload");
  if (!r)
    r = bhip(b, hipMemcpyAsync((void*)k.work, b->work.data(), sizeof(KWork) * b->work.size(),
                               hipMemcpyHostToDevice, b->stream), "tile list upload");
  // status words start clear (an upload, not a memset: nothing here is
  // captured, but the same holds for every buffer a graph reads)
  const std::vector<int32_t> zeros((size_t)p->n_design, 0);
  if (!r)
    r = bhip(b, hipMemcpyAsync(k.status, zeros.data(), sizeof(int32_t) * zeros.size(),
                               hipMemcpyHostToDevice, b->stream), "status upload");
  // the uploads read pageable host memory owned by this object: finish them
  if (!r) r = bhip(b, hipStreamSynchronize(b->stream), "upload sync");
  if (r) return r;
  b->step = p->table_step0;
  b->parity = 0;
  b->pend0 = b->step;
  b->pend_n = 0;
  b->configured = true;
  return 0;
}

int pm_batch_run(pm_batch* b, int32_t nsteps, uint32_t flags) {
  if (!b) return PM_EINVAL;
  if (!b->configured) return bfail(b, PM_ENOPLAN, "pm_batch_run before pm_batch_configure");
  if (nsteps < 0) return bfail(b, PM_EINVAL, "nsteps < 0");
  const pm_batch_plan& p = b->plan;
  if (b->step + nsteps > p.table_step0 + p.table_len)
    return bfail(b, PM_ERANGE, "source/phasor tables end at step %lld",
                 (long long)(p.table_step0 + p.table_len));
  (void)hipSetDevice(b->device);
  const int T = p.sample_T;
  const bool graphs = (flags & PM_RUN_GRAPH) && (T % 2 == 0);
  if ((flags & PM_RUN_GRAPH) && !graphs)
    return bfail(b, PM_EINVAL, "graph replay needs an even sample_T");
  if (graphs && b->stream == nullptr)
    return bfail(b, PM_EINVAL, "graph replay needs a non-default stream");
  int64_t left = nsteps;
  bool clock_ok = false;
  while (left > 0) {
    if (graphs && b->step % T == 0 && b->pend_n == 0 && left >= T) {
      int r = 0;
      if (!b->graph[b->parity]) r = capture(b, b->parity);
      if (r) return r;
      if (!clock_ok) {
        hipLaunchKernelGGL(k_clock_set, dim3(1), dim3(64), 0, b->stream, b->kb.clock, b->step);
        r = bhip(b, hipGetLastError(), "clock launch");
        if (r) return r;
        clock_ok = true;
      }
      {
        BSpan sp(b, 2);
        r = bhip(b, hipGraphLaunch(b->graph[b->parity], b->stream), "graph launch");
      }
      if (r) return r;
      b->step += T;
      b->pend0 = b->step;
      left -= T;
      continue;
    }
    clock_ok = false;
    int r;
    {
      BSpan sp(b, 0);
      r = step_once(b, b->step, 0, 0, b->parity);
    }
    if (r) return r;
    if (b->pend_n == 0) b->pend0 = b->step;
    ++b->pend_n;
    b->parity ^= 1;
    ++b->step;
    --left;
    if (b->step % T == 0) {
      BSpan sp(b, 1);
      r = flush_block(b, b->pend0, b->pend_n, nullptr);
      if (r) return r;
      b->pend_n = 0;
      b->pend0 = b->step;
    }
  }
  return 0;
}

int pm_batch_flush(pm_batch* b) {
  if (!b) return PM_EINVAL;
  if (!b->configured) return bfail(b, PM_ENOPLAN, "pm_batch_flush before configure");
  (void)hipSetDevice(b->device);
  int r = flush_block(b, b->pend0, b->pend_n, nullptr);
  if (r) return r;
  b->pend_n = 0;
  b->pend0 = b->step;
  return 0;
}

int pm_batch_parity(pm_batch* b) { return b ? b->parity : PM_EINVAL; }
int64_t pm_batch_steps(pm_batch* b) { return b ? b->step - b->plan.table_step0 : PM_EINVAL; }

int pm_batch_sync(pm_batch* b) {
  if (!b) return PM_EINVAL;
  (void)hipSetDevice(b->device);
  return bhip(b, hipStreamSynchronize(b->stream), "sync");
}

int pm_batch_set_tuning(pm_batch* b, int waves, int xcd_order) {
  if (!b) return PM_EINVAL;
  if (waves != 1 && waves != 2 && waves != 4 && waves != 8)
    return bfail(b, PM_EINVAL, "waves per workgroup must be 1, 2, 4 or 8");
  if (xcd_order != 0 && xcd_order != 1) return bfail(b, PM_EINVAL, "xcd_order must be 0 or 1");
  (void)hipSetDevice(b->device);
  // a captured graph holds the old launch shape
  drop_graphs(b);
  b->waves = b->kb.waves = waves;
  b->xcd = b->kb.xcd = xcd_order;
  return 0;
}

int pm_batch_design_status(pm_batch* b, int32_t* out, int32_t n) {
  if (!b || !out) return PM_EINVAL;
  if (!b->configured) return bfail(b, PM_ENOPLAN, "pm_batch_design_status before configure");
  if (n != b->plan.n_design)
    return bfail(b, PM_EINVAL, "status of %d designs asked, the plan has %d", n,
                 b->plan.n_design);
  (void)hipSetDevice(b->device);
  int r = bhip(b, hipStreamSynchronize(b->stream), "status sync");
  if (!r)
    r = bhip(b, hipMemcpyAsync(out, b->kb.status, sizeof(int32_t) * (size_t)n,
                               hipMemcpyDeviceToHost, b->stream), "status download");
  if (!r) r = bhip(b, hipStreamSynchronize(b->stream), "status sync");
  return r;
}

int pm_batch_set_timing(pm_batch* b, int enable) {
  if (!b) return PM_EINVAL;
  b->timing = enable != 0;
  for (auto& v : b->spans) {
    for (auto& ab : v) {
      (void)hipEventDestroy(ab.first);
      (void)hipEventDestroy(ab.second);
    }
    v.clear();
  }
  return 0;
}

int pm_batch_get_timing(pm_batch* b, int which, double* avg_ms, int64_t* count) {
  if (!b || which < 0 || which > 2 || !avg_ms || !count) return PM_EINVAL;
  (void)hipSetDevice(b->device);
  double tot = 0;
  int64_t n = 0;
  for (auto& ab : b->spans[which]) {
    float ms = 0;
    if (hipEventSynchronize(ab.second) != hipSuccess) continue;
    if (hipEventElapsedTime(&ms, ab.first, ab.second) == hipSuccess) {
      tot += ms;
      ++n;
    }
  }
  *avg_ms = n ? tot / n : 0.0;
  *count = n;
  return 0;
}

}  // extern "C"
