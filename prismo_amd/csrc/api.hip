// C ABI of libpmfdtd: context, plan validation, step loop, timing.
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "pm_internal.h"

using namespace pm;

struct pm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool configured = false;
  pm_plan plan{};
  KGrid g{};
  KAux aux{};
  std::vector<KDft> dft;
  std::string err;
  // spectral bookkeeping: steps [pend0, pend0 + pend_n) are sampled, not yet summed
  int64_t pend0 = 0;
  int32_t pend_n = 0;
  // fused ping-pong step: parity 0 = state in plan.E/H/psi_h, 1 = in *2
  bool fused_ok = false;
  int fuse = 1, tune_ty = 4, tune_kchunk = 8, tune_vec = 0, tune_order = 2;
  // rows in two segments: allowed for this plan (Cb-only coefficients, the
  // cell right of the first segment outside the x CPML layers), and the
  // tuning (0 auto, 1 whole rows, 2 two segments)
  bool split_ok = false;
  int tune_xs = 0;
  bool kchunk_set = false;  // pm_set_tuning was called (the 2-D pass's rows)
  // 2-D compat gate flags: the set the next H update fills (0 / 1) and the
  // set the E update of the current step reads
  int flag_set = 0, flag_set_e = 0;
  int parity = 0;
  // two steps per launch: plan eligible, and which E-side CPML memory set
  // holds the state (0 = plan.psi_e, 1 = plan.psi_e2)
  bool fused2_ok = false;
  // second stream for the slab tiles of a two-step pass (default priority)
  hipStream_t s2 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  bool src_kind0 = true;  // every source has one value per step (kind 0)
  int pe_parity = 0;
  KAux aux2{};  // source/sample view of the second buffer set
  std::vector<KAde> ade_host;  // staging of the ADE table upload
  std::vector<KSbox> sbox_host;  // staging of the box table uploads
  std::vector<KSrc> src_host;
  // timing
  bool timing = false;
  int timing_stride = 1;  // time the kernels of every timing_stride-th step
  bool span_on = false;   // timing, and the current step is one of those
  std::vector<hipEvent_t> pool;
  size_t pool_used = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> spans[7];
  // graph replay (PM_RUN_GRAPH): one captured block of steps per starting
  // buffer set (fused steps) or gate flag set (2-D compat steps), with the
  // run flags it was captured for
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  uint32_t graph_key[2] = {0, 0};
};

// Scratch of pm_configure_static: one 4 KiB copy on every device the
// library runs on (a __device__ array exists once per device), held until
// the process ends.  Host code never passes the array's name to the device:
// pm_create resolves each device's own address once, with that device
// current, and every later use takes that address.
static __device__ __align__(256) unsigned char scratch[PM_SCRATCH_BYTES];

namespace {

int fail(pm_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

int hip_check(pm_ctx* c, hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  return fail(c, (int)e, "%s: %s", what, hipGetErrorString(e));
}

void drop_graphs(pm_ctx* c) {
  for (auto& g : c->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// Timing events skip the system-scope fence on record: they only time the
// stream's own work (elapsed times are read after a stream synchronize), and
// the fence's cache writeback would otherwise add idle time between kernels.
hipEvent_t take_event(pm_ctx* c) {
  if (c->pool_used == c->pool.size()) {
    hipEvent_t e;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    c->pool.push_back(e);
  }
  return c->pool[c->pool_used++];
}

struct Span {
  pm_ctx* c;
  int which;
  hipEvent_t a = nullptr;
  Span(pm_ctx* ctx, int w) : c(ctx), which(w) {
    if (c->span_on) {
      a = take_event(c);
      if (a) (void)hipEventRecord(a, c->stream);
    }
  }
  ~Span() {
    if (a) {
      hipEvent_t b = take_event(c);
      if (b) {
        (void)hipEventRecord(b, c->stream);
        c->spans[which].push_back({a, b});
      }
    }
  }
};

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
int g_scratch_err[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Resolve device `dev`'s copy of `scratch` (dev must be current).
int resolve_static_scratch(int dev) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] || g_scratch_err[dev]) return g_scratch_err[dev];
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    g_scratch_err[dev] = e != hipSuccess ? (int)e : (int)hipErrorInvalidSymbol;
    return g_scratch_err[dev];
  }
  g_scratch_ptr[dev] = ptr;
  return 0;
}

bool box_ok(int i0, int i1, int j0, int j1, int k0, int k1, const pm_plan& p) {
  return i0 >= 0 && j0 >= 0 && k0 >= 0 && i0 <= i1 && j0 <= j1 && k0 <= k1 &&
         i1 <= p.nx && j1 <= p.ny && k1 <= p.nz;
}

// Kernel form of one dispersive box.  The box carries the reference's
// coefficients (src/prismo/materials/dispersion.py:189-231, 267-286); used
// as they stand (P+ = c0 E+ + c1 E + ..., J+ = c0 E+ + c1 J) the drive term
// sits half a step after the P/J recursion it feeds, which shifts the phase
// of chi(w) by w dt / 2 and shows up as a spurious gain Im(chi) ~ -Re(chi)
// w dt / 2.  The kernels average the drive over the two half steps instead:
//   Lorentz: P+ = c2 P + c3 P- + [(c0 E+ + c1 E) + (c0 E + c1 E-)] / 2
//            carried as P+ = a0 E+ + a1 E + a2 P + Q,  Q+ = a3 P + a4 E
//   Drude:   J+ = c1 J + c0 (E+ + E) / 2,  P+ = P + dt (J+ + J) / 2
// which keeps the static limit and leaves an O((w dt)^2) error against the
// closed-form eps(w).  E+ follows from eps_inf (E+ - E) + (P+ - P) = s with
// s = cb_scale curlH, P and J in units of eps0.
KAde centred_ade(const pm_ade_box& a, double dt) {
  KAde k{};
  k.comp = a.comp;
  k.kind = a.kind;
  k.i0 = a.i0; k.i1 = a.i1; k.j0 = a.j0; k.j1 = a.j1; k.k0 = a.k0; k.k1 = a.k1;
  k.p = a.p;
  const double einf = a.eps_inf;
  if (a.kind == 1) {
    k.q = a.p_prev;
    k.a0 = 0.5 * a.c0;
    k.a1 = 0.5 * (a.c0 + a.c1);
    k.a2 = a.c2;
    k.a3 = a.c3;
    k.a4 = 0.5 * a.c1;
    const double den = einf + k.a0;
    k.ke = (einf - k.a1) / den;
    k.kp = (1.0 - k.a2) / den;
    k.kq = 1.0 / den;
    k.ks = a.cb_scale / den;
  } else {
    const double h = 0.5 * dt;
    k.q = nullptr;
    k.a0 = 0.5 * a.c0;
    k.a1 = 0.5 * a.c0;
    k.a2 = a.c1;
    const double den = einf + h * k.a0;
    k.ke = (einf - h * k.a0) / den;
    k.kp = -h * (1.0 + a.c1) / den;
    k.kq = 0.0;
    k.ks = a.cb_scale / den;
  }
  return k;
}

int validate(pm_ctx* c, const pm_plan& p) {
  if (p.abi_version != PM_ABI_VERSION)
    return fail(c, PM_EINVAL, "plan abi_version %d != library %d", p.abi_version,
                PM_ABI_VERSION);
  if (p.ndim != 2 && p.ndim != 3) return fail(c, PM_EINVAL, "ndim must be 2 or 3");
  if (p.nx < 2 || p.ny < 2 || p.nz < 1) return fail(c, PM_EINVAL, "grid too small");
  if (p.ndim == 2 && p.nz != 1) return fail(c, PM_EINVAL, "2-D plan needs nz == 1");
  if (p.ndim == 3 && p.nz < 2) return fail(c, PM_EINVAL, "3-D plan needs nz >= 2");
  if (p.sy < p.nx || p.sz < p.sy * p.ny)
    return fail(c, PM_EINVAL, "strides (%lld, %lld) do not hold a %dx%d plane",
                (long long)p.sy, (long long)p.sz, p.nx, p.ny);
  if (!(p.dx > 0 && p.dy > 0 && p.dt > 0) || (p.ndim == 3 && !(p.dz > 0)))
    return fail(c, PM_EINVAL, "spacings and dt must be positive");
  if (p.precision != PM_FP64 && p.precision != PM_FP32)
    return fail(c, PM_EINVAL, "unknown precision %d", p.precision);
  if (p.coeff_kind < PM_COEFF_FULL || p.coeff_kind > PM_COEFF_CB_F16)
    return fail(c, PM_EINVAL, "unknown coeff_kind %d", p.coeff_kind);
  if (p.coeff_kind == PM_COEFF_CB_F16 && p.precision != PM_FP32)
    return fail(c, PM_ENOTSUP, "FP16 coefficients need FP32 fields");

  // which components the scheme touches
  bool needE[3], needH[3];
  if (p.scheme == PM_SCHEME_COMPAT) {
    if (p.precision != PM_FP64)
      return fail(c, PM_ENOTSUP, "compat scheme is FP64 only");
    if (p.coeff_kind == PM_COEFF_CB_F16)
      return fail(c, PM_ENOTSUP, "compat scheme needs FP64 coefficients");
    for (int a = 0; a < 6; ++a)
      if (p.pml[a]) return fail(c, PM_ENOTSUP, "compat scheme has no CPML");
    if (p.n_ade) return fail(c, PM_ENOTSUP, "compat scheme has no ADE");
    for (int a = 0; a < 3; ++a) needE[a] = needH[a] = true;
    if (p.ndim == 2 && !p.scratch)
      return fail(c, PM_EINVAL, "2-D compat needs scratch for its gates");
  } else if (p.scheme == PM_SCHEME_YEE) {
    for (int a = 0; a < 3; ++a) needE[a] = needH[a] = (p.ndim == 3);
    needE[2] = needH[0] = needH[1] = true;
  } else {
    return fail(c, PM_EINVAL, "unknown scheme %d", p.scheme);
  }
  for (int a = 0; a < 3; ++a) {
    if (needE[a] && !p.E[a]) return fail(c, PM_EINVAL, "E[%d] is NULL", a);
    if (needH[a] && !p.H[a]) return fail(c, PM_EINVAL, "H[%d] is NULL", a);
    if (needE[a] && !p.cb[a]) return fail(c, PM_EINVAL, "cb[%d] is NULL", a);
    if (p.coeff_kind == PM_COEFF_FULL) {
      if (needE[a] && !p.ca[a]) return fail(c, PM_EINVAL, "ca[%d] is NULL", a);
      if (needH[a] && (!p.da[a] || !p.db[a]))
        return fail(c, PM_EINVAL, "da/db[%d] is NULL", a);
    }
  }
  // CPML
  const int n[3] = {p.nx, p.ny, p.nz};
  for (int a = 0; a < 3; ++a) {
    const int t = p.pml[2 * a] + p.pml[2 * a + 1];
    if (p.pml[2 * a] < 0 || p.pml[2 * a + 1] < 0 || t > n[a])
      return fail(c, PM_EINVAL, "bad CPML thickness on axis %d", a);
    if (t == 0) continue;
    if (a == 2 && p.ndim == 2) return fail(c, PM_EINVAL, "2-D plan with z CPML");
    if (!p.pml_be[a] || !p.pml_ce[a] || !p.pml_ike[a] || !p.pml_bh[a] || !p.pml_ch[a] ||
        !p.pml_ikh[a])
      return fail(c, PM_EINVAL, "CPML profiles missing on axis %d", a);
    // psi[a][q]: q = 1 for component (a+2)%3 differentiated first, q = 0 the other
    for (int q = 0; q < 2; ++q) {
      const int comp = q == 1 ? (a + 2) % 3 : (a + 1) % 3;
      if (needE[comp] && !p.psi_e[a][q])
        return fail(c, PM_EINVAL, "psi_e[%d][%d] missing", a, q);
      if (needH[comp] && !p.psi_h[a][q])
        return fail(c, PM_EINVAL, "psi_h[%d][%d] missing", a, q);
    }
  }
  if (p.n_ade < 0 || p.n_ade > kMaxAde) return fail(c, PM_EINVAL, "too many ADE boxes");
  if (!p.scratch) return fail(c, PM_EINVAL, "scratch (PM_SCRATCH_BYTES) is required");
  static_assert(kScratchAde + sizeof(KAde) * kMaxAde <= kScratchSbox, "scratch layout");
  static_assert(kScratchSbox + sizeof(KSbox) * kMaxSbox <= kScratchSrc, "scratch layout");
  static_assert(kScratchSrc + sizeof(KSrc) * kMaxSrc <= PM_SCRATCH_BYTES, "scratch layout");
  for (int b = 0; b < p.n_ade; ++b) {
    const pm_ade_box& a = p.ade[b];
    if (a.comp < 0 || a.comp > 2 || !needE[a.comp] || (a.kind != 1 && a.kind != 2))
      return fail(c, PM_EINVAL, "ADE box %d: bad component/kind", b);
    if (!box_ok(a.i0, a.i1, a.j0, a.j1, a.k0, a.k1, p) || !a.p || (a.kind == 1 && !a.p_prev))
      return fail(c, PM_EINVAL, "ADE box %d: bad extent or state", b);
    if (!(a.eps_inf > 0.0) || !(a.cb_scale > 0.0))
      return fail(c, PM_EINVAL, "ADE box %d: eps_inf and cb_scale must be positive", b);
  }
  if (p.n_src < 0 || p.n_src > kMaxSrc) return fail(c, PM_EINVAL, "too many sources");
  for (int s = 0; s < p.n_src; ++s) {
    const pm_source& q = p.src[s];
    if (q.comp < 0 || q.comp > 5) return fail(c, PM_EINVAL, "source %d: bad comp", s);
    const bool present = q.comp < 3 ? needE[q.comp] : needH[q.comp - 3];
    if (!present) return fail(c, PM_EINVAL, "source %d: component not simulated", s);
    if (!box_ok(q.i0, q.i1, q.j0, q.j1, q.k0, q.k1, p) || !q.table ||
        (q.kind == 1 && !q.profile) || (q.kind != 0 && q.kind != 1))
      return fail(c, PM_EINVAL, "source %d: bad box/table", s);
  }
  if (p.n_src && p.table_len <= 0) return fail(c, PM_EINVAL, "empty source table");
  if (p.n_sbox < 0 || p.n_sbox > kMaxSbox) return fail(c, PM_EINVAL, "too many monitor boxes");
  if (p.n_sbox && (!p.X || p.sample_T <= 0 || p.ncols <= 0))
    return fail(c, PM_EINVAL, "monitor boxes need X, sample_T and ncols");
  for (int s = 0; s < p.n_sbox; ++s) {
    const pm_sample_box& b = p.sbox[s];
    if (b.comp < 0 || b.comp > 5) return fail(c, PM_EINVAL, "monitor box %d: bad comp", s);
    const bool present = b.comp < 3 ? needE[b.comp] : needH[b.comp - 3];
    if (!present) return fail(c, PM_EINVAL, "monitor box %d: component not simulated", s);
    const int64_t n = (int64_t)(b.i1 - b.i0) * (b.j1 - b.j0) * (b.k1 - b.k0);
    if (!box_ok(b.i0, b.i1, b.j0, b.j1, b.k0, b.k1, p) || b.col0 < 0 || b.col0 + n > p.ncols)
      return fail(c, PM_EINVAL, "monitor box %d: bad extent/columns", s);
  }
  if (p.n_dft < 0 || p.n_dft > kMaxDft) return fail(c, PM_EINVAL, "too many DFT groups");
  for (int s = 0; s < p.n_dft; ++s) {
    const pm_dft_group& d = p.dft[s];
    if (d.col0 < 0 || d.col1 > p.ncols || d.col0 >= d.col1 || d.f0 < 0 ||
        d.f1 > p.nfreq_total || d.f0 >= d.f1 || !d.acc_re || !d.acc_im ||
        ((d.cmp_re == nullptr) != (d.cmp_im == nullptr)))
      return fail(c, PM_EINVAL, "DFT group %d: bad ranges/buffers", s);
    if (!p.W_re || !p.W_im || p.table_len <= 0)
      return fail(c, PM_EINVAL, "DFT groups need phasor tables");
  }
  return 0;
}

int launch_status(pm_ctx* c, const char* what) {
  return hip_check(c, hipGetLastError(), what);
}

// The gate flag sets of a 2-D compat step: `set` filled before the H
// update and read by it and the E update, the other one cleared by the H
// kernel for the next step.
KGrid flag_grid(const pm_ctx* c, int set) {
  KGrid g = c->g;
  g.flags = c->g.flags + 8 * set;
  g.flags_next = c->g.flags + 8 * (set ^ 1);
  return g;
}

int do_h(pm_ctx* c) {
  Span sp(c, 0);
  const pm_plan& p = c->plan;
  if (p.scheme == PM_SCHEME_COMPAT) {
    if (p.ndim == 2) {
      const KGrid g = flag_grid(c, c->flag_set);
      launch_compat_flags2d(g, c->stream);
      launch_compat_h(g, p.coeff_kind, c->stream);
      c->flag_set_e = c->flag_set;
      c->flag_set ^= 1;
    } else {
      launch_compat_h(c->g, p.coeff_kind, c->stream);
    }
  } else {
    launch_yee_h(c->g, p.precision, p.coeff_kind, c->stream);
  }
  return launch_status(c, "H update launch");
}

int do_e(pm_ctx* c) {
  Span sp(c, 1);
  const pm_plan& p = c->plan;
  if (p.scheme == PM_SCHEME_COMPAT)
    launch_compat_e(p.ndim == 2 ? flag_grid(c, c->flag_set_e) : c->g, p.coeff_kind, c->stream);
  else
    launch_yee_e(c->g, p.precision, p.coeff_kind, c->stream);
  return launch_status(c, "E update launch");
}

// Cells per lane and waves per row of the fused kernel for this plan.
// Launch shape of the fused kernel: the tuned one, else (rows per workgroup
// and then cells per lane relaxed) the first that fits the workgroup and LDS
// limits.  False if none does.
// Rows per tile of the 2-D pass: kchunk once pm_set_tuning has set it, else
// 3.  A wave's rows run one after another (two rows of loads in flight), so
// short tiles and many waves suit the small grids 2-D runs have: on a
// 540 x 290 FP64 grid 2 and 3 rows per tile step in 7.3 us, 8 rows in
// 11.5 us, the unfused H and E launches in 8.3 us.
int rows2d(const pm_ctx* c) { return c->kchunk_set ? std::max(1, c->tune_kchunk) : 3; }

// Rows in xs segments (pm_set_row_split: 1 or 2; 0 = whole rows when a
// shape fits them, else two segments).
bool fused_shape(const pm_ctx* c, int* vec, int* w, int* ty, int* xs) {
  const int sy = (int)c->plan.sy;  // padded row, 128 / 256 / 512
  if (c->plan.ndim == 2) {
    // the 2-D pass: one cell per lane, a wave per 64-cell strip, tiles of
    // rows2d rows
    *vec = 1;
    *w = sy / 64;
    *ty = rows2d(c);
    *xs = 1;
    return true;
  }
  const int prec = c->plan.precision, ck = c->plan.coeff_kind;
  int opts[2] = {1, 2}, nopt = 2;
  if (c->tune_xs == 1 || c->tune_xs == 2) {
    opts[0] = c->tune_xs;
    nopt = 1;
  }
  for (int a = 0; a < nopt; ++a) {
    const int x = opts[a];
    if (x == 2 && !c->split_ok) continue;
    const int rw = sy / x;  // cells per segment
    int prefer = c->tune_vec;
    if (prefer <= 0) prefer = prec == PM_FP32 ? (rw >= 128 && x == 1 ? 2 : 1) : (rw <= 256 ? 1 : 2);
    const int vecs[4] = {prefer, 1, 2, 4};
    const int nv = c->tune_vec > 0 ? 1 : 4;
    for (int iv = 0; iv < nv; ++iv) {
      const int v = vecs[iv];
      if (rw % (64 * v)) continue;
      for (int t : {c->tune_ty, 2}) {
        if (fused_supported(prec, ck, v, rw / (64 * v), t, x)) {
          *vec = v;
          *w = rw / (64 * v);
          *ty = t;
          *xs = x;
          return true;
        }
      }
    }
  }
  const int prefer = c->tune_vec > 0 ? c->tune_vec : 1;
  *vec = prefer;
  *w = sy / (64 * prefer);
  *ty = c->tune_ty;
  *xs = 1;
  return false;
}

FusedBufs fused_bufs(const pm_plan& p, bool a_in) {
  FusedBufs fb{};  // pse_out stays NULL: the one-step kernel updates psi_e in place
  for (int a = 0; a < 3; ++a) {
    fb.Ei[a] = a_in ? p.E[a] : p.E2[a];
    fb.Hi[a] = a_in ? p.H[a] : p.H2[a];
    fb.Eo[a] = a_in ? p.E2[a] : p.E[a];
    fb.Ho[a] = a_in ? p.H2[a] : p.H[a];
    for (int q = 0; q < 2; ++q) {
      fb.psh_in[a][q] = a_in ? p.psi_h[a][q] : p.psi_h2[a][q];
      fb.psh_out[a][q] = a_in ? p.psi_h2[a][q] : p.psi_h[a][q];
    }
  }
  return fb;
}

// One fused step from buffer set `parity` into the other one.
int do_fused(pm_ctx* c) {
  Span sp(c, 3);
  const pm_plan& p = c->plan;
  if (p.ndim == 2) {
    const bool a_in = c->parity == 0;
    Tm2dBufs b{};
    b.Ei = a_in ? p.E[2] : p.E2[2];
    b.Eo = a_in ? p.E2[2] : p.E[2];
    for (int q = 0; q < 2; ++q) {
      b.Hi[q] = a_in ? p.H[q] : p.H2[q];
      b.Ho[q] = a_in ? p.H2[q] : p.H[q];
    }
    // H-side CPML memory: Hy in the x slabs (axis 0, first array), Hx in
    // the y slabs (axis 1, second array)
    b.psh_in[0] = a_in ? p.psi_h[0][0] : p.psi_h2[0][0];
    b.psh_out[0] = a_in ? p.psi_h2[0][0] : p.psi_h[0][0];
    b.psh_in[1] = a_in ? p.psi_h[1][1] : p.psi_h2[1][1];
    b.psh_out[1] = a_in ? p.psi_h2[1][1] : p.psi_h[1][1];
    if (launch_tm2d_fused(c->g, p.precision, p.coeff_kind, b, rows2d(c), c->stream))
      return fail(c, PM_ENOTSUP, "fused 2-D step: plan not supported");
    c->parity ^= 1;
    return launch_status(c, "fused 2-D step launch");
  }
  FusedBufs fb{};
  const bool a_in = c->parity == 0;
  for (int a = 0; a < 3; ++a) {
    fb.Ei[a] = a_in ? p.E[a] : p.E2[a];
    fb.Hi[a] = a_in ? p.H[a] : p.H2[a];
    fb.Eo[a] = a_in ? p.E2[a] : p.E[a];
    fb.Ho[a] = a_in ? p.H2[a] : p.H[a];
    for (int q = 0; q < 2; ++q) {
      fb.psh_in[a][q] = a_in ? p.psi_h[a][q] : p.psi_h2[a][q];
      fb.psh_out[a][q] = a_in ? p.psi_h2[a][q] : p.psi_h[a][q];
    }
  }
  int vec, w, ty, xs;
  fused_shape(c, &vec, &w, &ty, &xs);
  const int r = p.precision == PM_FP32
                    ? launch_yee_fused_f32(c->g, p.coeff_kind, fb, vec, w, ty,
                                           c->tune_kchunk, c->tune_order, xs, c->stream)
                    : launch_yee_fused_f64(c->g, p.coeff_kind, fb, vec, w, ty,
                                           c->tune_kchunk, c->tune_order, xs, c->stream);
  if (r != 0)
    return fail(c, PM_ENOTSUP, "fused step: no kernel for vec %d, w %d, ty %d, %d segments",
                vec, w, ty, xs);
  c->parity ^= 1;
  return launch_status(c, "fused step launch");
}

// Point the kernels' E-side CPML memory at the set that holds the state.
void set_psi_e(pm_ctx* c) {
  for (int a = 0; a < 3; ++a)
    for (int q = 0; q < 2; ++q)
      c->g.psi_e[a][q] = c->pe_parity ? c->plan.psi_e2[a][q] : c->plan.psi_e[a][q];
}

// Launch shape of the two-step kernel: the padded row decides the cells per
// lane (two waves per row).
bool fused2_shape(const pm_ctx* c, int* vec, int* w) {
  const int rw = (int)c->plan.sy;
  *w = 2;
  *vec = rw / 128;
  return (rw == 128 || rw == 256) && fused2_supported(*vec, *w, 4);
}

FusedBufs fused2_bufs(const pm_ctx* c) {
  const pm_plan& p = c->plan;
  FusedBufs fb = fused_bufs(p, c->parity == 0);
  for (int a = 0; a < 3; ++a)
    for (int q = 0; q < 2; ++q) fb.pse_out[a][q] = c->pe_parity ? p.psi_e[a][q] : p.psi_e2[a][q];
  return fb;
}

// Two steps (n, n+1) in one launch from the current sets into the others,
// with the sources of step n added inside the kernel.
int do_fused2(pm_ctx* c, int64_t n, bool src) {
  Span sp(c, 5);
  const pm_plan& p = c->plan;
  int vec, w;
  fused2_shape(c, &vec, &w);
  const FusedBufs fb = fused2_bufs(c);
  const KAux& ax = c->parity ? c->aux2 : c->aux;
  const int r = launch_yee_fused2_f32(c->g, p.coeff_kind, fb, vec, w, 4, c->tune_kchunk,
                                      c->tune_order, src ? n - p.table_step0 : 0,
                                      src ? ax.n_src : 0, ax.src_dev, c->stream, c->s2,
                                      c->fork, c->join);
  if (r == -2) return hip_check(c, hipGetLastError(), "two-step stream fork/join");
  if (r != 0) return fail(c, PM_ENOTSUP, "two-step launch: no kernel for vec %d, w %d", vec, w);
  c->parity ^= 1;
  c->pe_parity ^= 1;
  set_psi_e(c);
  return launch_status(c, "two-step launch");
}

bool fused2_eligible(const pm_ctx* c) {
  const pm_plan& p = c->plan;
  if (!c->fused_ok || p.ndim != 3 || p.precision != PM_FP32 || p.n_ade ||
      !fused2_grid_ok(c->g))
    return false;
  if (p.coeff_kind != PM_COEFF_CB && p.coeff_kind != PM_COEFF_CB_F16) return false;
  int vec, w;
  if (!fused2_shape(c, &vec, &w)) return false;
  for (int a = 0; a < 3; ++a)
    for (int q = 0; q < 2; ++q)
      if (p.psi_e[a][q] && !p.psi_e2[a][q]) return false;
  for (int pp = 0; pp < 2; ++pp) {
    for (int pe = 0; pe < 2; ++pe) {
      KGrid g = c->g;
      FusedBufs fb = fused_bufs(p, pp == 0);
      for (int a = 0; a < 3; ++a)
        for (int q = 0; q < 2; ++q) {
          g.psi_e[a][q] = pe ? p.psi_e2[a][q] : p.psi_e[a][q];
          fb.pse_out[a][q] = pe ? p.psi_e[a][q] : p.psi_e2[a][q];
        }
      if (!fused_addressable(g, fb, 4)) return false;
    }
  }
  return true;
}

// The 2-D TM one-pass step (tm2d.hip): Cb-only coefficients, no dispersive
// boxes, a second set for Ez, Hx, Hy and for the H-side CPML memory.
bool fused2d_eligible(const pm_plan& p) {
  if (p.scheme != PM_SCHEME_YEE || p.ndim != 2 || p.n_ade) return false;
  if (p.coeff_kind != PM_COEFF_CB &&
      !(p.coeff_kind == PM_COEFF_CB_F16 && p.precision == PM_FP32))
    return false;
  if (!p.E2[2] || !p.H2[0] || !p.H2[1]) return false;
  if ((uint64_t)p.sy * (uint64_t)p.ny >= (1ull << 31)) return false;
  for (int a = 0; a < 2; ++a)
    for (int q = 0; q < 2; ++q)
      if (p.psi_h[a][q] && !p.psi_h2[a][q]) return false;
  return true;
}

bool fused_eligible(const pm_plan& p) {
  if (p.ndim == 2) return fused2d_eligible(p);
  if (p.scheme != PM_SCHEME_YEE || p.ndim != 3) return false;
  if (p.sy != 128 && p.sy != 256 && p.sy != 512) return false;
  if (p.nx > p.sy) return false;
  // the fused kernel addresses fields with 32-bit byte offsets
  const uint64_t esz = p.precision == PM_FP64 ? 8 : 4;
  if ((uint64_t)p.sz * (uint64_t)p.nz * esz >= (1ull << 32)) return false;
  for (int a = 0; a < 3; ++a) {
    if (!p.E2[a] || !p.H2[a]) return false;
    for (int q = 0; q < 2; ++q)
      if (p.psi_h[a][q] && !p.psi_h2[a][q]) return false;
  }
  return true;
}

// Sum rows (first mod T) .. (first mod T) + rows - 1 of X, the samples of
// steps [first, first + rows), into every DFT group.
int flush_rows(pm_ctx* c, int64_t first, int32_t rows) {
  const pm_plan& p = c->plan;
  if (rows <= 0) return 0;
  const int64_t wcol0 = first - p.table_step0;
  if (wcol0 < 0 || wcol0 + rows > p.table_len)
    return fail(c, PM_ERANGE, "phasor table does not cover steps %lld..%lld",
                (long long)first, (long long)(first + rows - 1));
  const int esz = p.precision == PM_FP64 ? 8 : 4;
  const char* X0 = (const char*)p.X + (first % p.sample_T) * (int64_t)p.ncols * esz;
  Span sp(c, 4);
  for (const KDft& d : c->dft)
    launch_dft(d, X0, p.ncols, p.W_re, p.W_im, p.table_len, wcol0, rows, p.precision,
               c->stream);
  return launch_status(c, "DFT launch");
}

int flush_pending(pm_ctx* c) {
  if (c->pend_n == 0) return 0;
  const int r = flush_rows(c, c->pend0, c->pend_n);
  if (r == 0) c->pend_n = 0;
  return r;
}

}  // namespace

extern "C" {

int pm_abi_version(void) { return PM_ABI_VERSION; }

int pm_create(int device, void* stream, pm_ctx** out) {
  if (!out) return PM_EINVAL;
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess) return (int)e;
  if (device < 0 || device >= ndev) return PM_EINVAL;
  e = hipSetDevice(device);
  if (e != hipSuccess) return (int)e;
  // the device's copy of the static scratch (pm_configure_static); a
  // failure here only disables that entry point
  if (device < kMaxDevices) (void)resolve_static_scratch(device);
  pm_ctx* c = new pm_ctx();
  c->device = device;
  c->stream = (hipStream_t)stream;
  // the two-step pass's second stream and its fork/join events; without
  // them the pass runs its two launches one after the other
  if (hipStreamCreateWithFlags(&c->s2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->join, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->join) (void)hipEventDestroy(c->join);
    if (c->s2) (void)hipStreamDestroy(c->s2);
    c->s2 = nullptr;
    c->fork = c->join = nullptr;
  }
  *out = c;
  return 0;
}

int pm_destroy(pm_ctx* c) {
  if (!c) return 0;
  if (c->s2) (void)hipStreamSynchronize(c->s2);
  drop_graphs(c);
  if (c->fork) (void)hipEventDestroy(c->fork);
  if (c->join) (void)hipEventDestroy(c->join);
  if (c->s2) (void)hipStreamDestroy(c->s2);
  for (hipEvent_t e : c->pool) (void)hipEventDestroy(e);
  delete c;
  return 0;
}

const char* pm_last_error(pm_ctx* c) { return c ? c->err.c_str() : "null context"; }

int pm_configure(pm_ctx* c, const pm_plan* plan) {
  if (!c || !plan) return PM_EINVAL;
  c->configured = false;
  // captured graphs hold the old plan's pointers; a replay may still run
  if (c->s2) (void)hipStreamSynchronize(c->s2);
  drop_graphs(c);
  const int v = validate(c, *plan);
  if (v) return v;
  const pm_plan& p = *plan;
  c->plan = p;
  // detach the host arrays
  c->plan.ade = nullptr;
  c->plan.src = nullptr;
  c->plan.sbox = nullptr;
  c->plan.dft = nullptr;

  KGrid& g = c->g;
  g = KGrid{};
  g.nx = p.nx;
  g.ny = p.ny;
  g.nz = p.nz;
  g.ndim = p.ndim;
  g.sy = p.sy;
  g.sz = p.sz;
  g.dx = p.dx;
  g.dy = p.dy;
  g.dz = p.dz;
  g.dt = p.dt;
  g.idx = 1.0 / p.dx;
  g.idy = 1.0 / p.dy;
  g.idz = p.ndim == 3 ? 1.0 / p.dz : 0.0;
  for (int a = 0; a < 3; ++a) {
    g.E[a] = p.E[a];
    g.H[a] = p.H[a];
    g.ca[a] = p.ca[a];
    g.cb[a] = p.cb[a];
    g.da[a] = p.da[a];
    g.db[a] = p.db[a];
    g.pml_be[a] = p.pml_be[a];
    g.pml_ce[a] = p.pml_ce[a];
    g.pml_ike[a] = p.pml_ike[a];
    g.pml_bh[a] = p.pml_bh[a];
    g.pml_ch[a] = p.pml_ch[a];
    g.pml_ikh[a] = p.pml_ikh[a];
    for (int q = 0; q < 2; ++q) {
      g.psi_e[a][q] = p.psi_e[a][q];
      g.psi_h[a][q] = p.psi_h[a][q];
    }
  }
  g.cb_scale = p.cb_scale;
  g.db_const = p.db_const;
  g.has_pml = 0;
  for (int a = 0; a < 6; ++a) {
    g.pml[a] = p.pml[a];
    if (p.pml[a]) g.has_pml = 1;
  }
  g.n_ade = p.n_ade;
  g.flags = (int*)((char*)p.scratch + kScratchFlags);
  g.ade = (const KAde*)((char*)p.scratch + kScratchAde);
  g.ade_j0 = g.ade_k0 = 1 << 30;
  g.ade_j1 = g.ade_k1 = -1;
  for (int b = 0; b < p.n_ade; ++b) {
    g.ade_j0 = std::min(g.ade_j0, p.ade[b].j0);
    g.ade_j1 = std::max(g.ade_j1, p.ade[b].j1);
    g.ade_k0 = std::min(g.ade_k0, p.ade[b].k0);
    g.ade_k1 = std::max(g.ade_k1, p.ade[b].k1);
  }
  if (p.n_ade) {
    KAde tab[kMaxAde];
    for (int b = 0; b < p.n_ade; ++b) {
      const pm_ade_box& a = p.ade[b];
      tab[b] = centred_ade(a, p.dt);
    }
    // stream-ordered upload into caller memory; the host copy lives in ctx
    c->ade_host.assign(tab, tab + p.n_ade);
    int r = hip_check(c, hipMemcpyAsync((void*)g.ade, c->ade_host.data(),
                                        sizeof(KAde) * p.n_ade, hipMemcpyHostToDevice,
                                        c->stream), "ADE table upload");
    if (r) return r;
  }

  KAux& x = c->aux;
  x = KAux{};
  x.nx = p.nx;
  x.ny = p.ny;
  x.nz = p.nz;
  x.sy = p.sy;
  x.sz = p.sz;
  for (int a = 0; a < 3; ++a) {
    x.F[a] = p.E[a];
    x.F[3 + a] = p.H[a];
  }
  x.precision = p.precision;
  x.n_src = p.n_src;
  for (int s = 0; s < p.n_src; ++s) {
    const pm_source& q = p.src[s];
    x.src[s] = KSrc{q.comp, q.kind, q.i0, q.i1, q.j0, q.j1, q.k0, q.k1,
                    q.table, q.profile, q.divisor};
  }
  x.table_step0 = p.table_step0;
  x.table_len = p.table_len;
  x.n_sbox = p.n_sbox;
  x.sample_T = p.sample_T;
  x.ncols = p.ncols;
  x.X = p.X;
  for (int s = 0; s < p.n_sbox; ++s) {
    const pm_sample_box& b = p.sbox[s];
    x.sbox[s] = KSbox{b.comp, b.i0, b.i1, b.j0, b.j1, b.k0, b.k1, b.col0};
  }
  // box tables for the merged launches, uploaded (stream-ordered) into the
  // caller's scratch; the host copies live in ctx until the next configure
  auto cells = [](int i0, int i1, int j0, int j1, int k0, int k1) {
    return (int64_t)std::max(i1 - i0, 0) * std::max(j1 - j0, 0) * std::max(k1 - k0, 0);
  };
  x.sbox_dev = (const KSbox*)((char*)p.scratch + kScratchSbox);
  x.src_dev = (const KSrc*)((char*)p.scratch + kScratchSrc);
  x.sbox_max = x.src_max = 0;
  for (int s = 0; s < x.n_sbox; ++s) {
    const KSbox& b = x.sbox[s];
    x.sbox_max = std::max(x.sbox_max, cells(b.i0, b.i1, b.j0, b.j1, b.k0, b.k1));
  }
  x.src_merge = true;
  for (int s = 0; s < x.n_src; ++s) {
    const KSrc& a = x.src[s];
    x.src_max = std::max(x.src_max, cells(a.i0, a.i1, a.j0, a.j1, a.k0, a.k1));
    for (int t = 0; t < s; ++t) {
      const KSrc& b = x.src[t];
      if (a.comp == b.comp && a.i0 < b.i1 && b.i0 < a.i1 && a.j0 < b.j1 && b.j0 < a.j1 &&
          a.k0 < b.k1 && b.k0 < a.k1)
        x.src_merge = false;
    }
  }
  x.io_merge = x.src_merge && x.n_src > 0 && x.n_sbox > 0 && x.src_max > 0 && x.sbox_max > 0;
  for (int s = 0; s < x.n_sbox && x.io_merge; ++s)
    for (int t = 0; t < x.n_src; ++t) {
      const KSbox& b = x.sbox[s];
      const KSrc& a = x.src[t];
      if (a.comp == b.comp && a.i0 < b.i1 && b.i0 < a.i1 && a.j0 < b.j1 && b.j0 < a.j1 &&
          a.k0 < b.k1 && b.k0 < a.k1)
        x.io_merge = false;
    }
  c->sbox_host.assign(x.sbox, x.sbox + x.n_sbox);
  c->src_host.assign(x.src, x.src + x.n_src);
  if (x.n_sbox) {
    int r = hip_check(c, hipMemcpyAsync((void*)x.sbox_dev, c->sbox_host.data(),
                                        sizeof(KSbox) * x.n_sbox, hipMemcpyHostToDevice,
                                        c->stream), "sample table upload");
    if (r) return r;
  }
  if (x.n_src) {
    int r = hip_check(c, hipMemcpyAsync((void*)x.src_dev, c->src_host.data(),
                                        sizeof(KSrc) * x.n_src, hipMemcpyHostToDevice,
                                        c->stream), "source table upload");
    if (r) return r;
  }
  c->dft.clear();
  for (int s = 0; s < p.n_dft; ++s) {
    const pm_dft_group& d = p.dft[s];
    c->dft.push_back(KDft{d.col0, d.col1, d.f0, d.f1, d.acc_re, d.acc_im, d.cmp_re, d.cmp_im});
  }
  c->aux2 = x;
  for (int a = 0; a < 3; ++a) {
    c->aux2.F[a] = p.E2[a];
    c->aux2.F[3 + a] = p.H2[a];
  }
  // the uploads read host staging that the next configure overwrites:
  // wait for them here (configure is set-up, not the step loop)
  if (p.n_ade || x.n_sbox || x.n_src) {
    int r = hip_check(c, hipStreamSynchronize(c->stream), "table upload sync");
    if (r) return r;
  }
  {
    // two row segments: the cell right of the first one (x = sy / 2) must
    // lie inside the grid and outside the x CPML layers
    const int gxr = (int)(p.sy / 2);
    c->split_ok = p.coeff_kind != PM_COEFF_FULL && gxr < p.nx && gxr >= p.pml[0] &&
                  gxr < p.nx - p.pml[1];
  }
  c->fused_ok = p.ndim == 2
                    ? fused2d_eligible(p)
                    : fused_eligible(p) &&
                          fused_addressable(c->g, fused_bufs(p, true), p.precision == PM_FP64 ? 8 : 4) &&
                          fused_addressable(c->g, fused_bufs(p, false), p.precision == PM_FP64 ? 8 : 4);
  c->parity = 0;
  c->pe_parity = 0;
  set_psi_e(c);
  c->fused2_ok = fused2_eligible(c);
  c->src_kind0 = true;
  for (int s = 0; s < x.n_src; ++s)
    if (x.src[s].kind != 0) c->src_kind0 = false;
  c->pend_n = 0;
  // both gate flag sets start cleared (2-D compat)
  c->flag_set = c->flag_set_e = 0;
  if (p.scheme == PM_SCHEME_COMPAT && p.ndim == 2) {
    launch_compat_flags_clear(c->g.flags, c->stream);
    int r = launch_status(c, "flag clear launch");
    if (r) return r;
  }
  c->configured = true;
  return 0;
}

int pm_configure_static(pm_ctx* c, const pm_plan* plan) {
  if (!c || !plan) return PM_EINVAL;
  if (plan->scratch)
    return fail(c, PM_EINVAL,
                "pm_configure_static: plan.scratch must be NULL (the library's own "
                "%d-byte scratch is used)", PM_SCRATCH_BYTES);
  // what the scratch holds, against its fixed size (validate repeats these)
  if (plan->n_ade > kMaxAde || plan->n_sbox > kMaxSbox || plan->n_src > kMaxSrc)
    return fail(c, PM_EINVAL,
                "pm_configure_static: %d dispersive boxes, %d monitor boxes, %d sources "
                "exceed the %d-byte static scratch (at most %d, %d, %d)",
                plan->n_ade, plan->n_sbox, plan->n_src, PM_SCRATCH_BYTES, kMaxAde, kMaxSbox,
                kMaxSrc);
  void* ptr = c->device < kMaxDevices ? g_scratch_ptr[c->device] : nullptr;
  if (!ptr)
    return fail(c, PM_ENOTSUP,
                "pm_configure_static: no static scratch on device %d (symbol lookup status %d)",
                c->device, c->device < kMaxDevices ? g_scratch_err[c->device] : -1);
  pm_plan p = *plan;
  p.scratch = ptr;
  return pm_configure(c, &p);
}

int pm_update_h(pm_ctx* c) {
  if (!c) return PM_EINVAL;
  if (!c->configured) return fail(c, PM_ENOPLAN, "pm_update_h before pm_configure");
  return do_h(c);
}

int pm_update_e(pm_ctx* c) {
  if (!c) return PM_EINVAL;
  if (!c->configured) return fail(c, PM_ENOPLAN, "pm_update_e before pm_configure");
  return do_e(c);
}

namespace {

// Sample step n from the current state (after its sources) and sum a
// completed block; a gap in the sampled steps sums (or drops) the block.
// Before step n is sampled: samples pending from a block that does not
// continue at n are summed (or dropped without spectra).
int sample_pre(pm_ctx* c, int64_t n, bool dft) {
  if (c->pend_n && n != c->pend0 + c->pend_n) {
    if (dft) return flush_pending(c);
    c->pend_n = 0;
  }
  return 0;
}

// After step n is sampled: it joins the pending block, which is summed
// when it completes.
int sample_post(pm_ctx* c, int64_t n, bool dft);

int sample_step(pm_ctx* c, int64_t n, bool dft) {
  int r;
  if ((r = sample_pre(c, n, dft))) return r;
  const KAux& ax = c->parity ? c->aux2 : c->aux;
  launch_sample(ax, n, c->stream);
  if ((r = launch_status(c, "sample launch"))) return r;
  return sample_post(c, n, dft);
}

int sample_post(pm_ctx* c, int64_t n, bool dft) {
  const pm_plan& p = c->plan;
  int r;
  if (c->pend_n == 0) c->pend0 = n;
  ++c->pend_n;
  if ((n + 1) % p.sample_T == 0) {
    if (dft) {
      if ((r = flush_pending(c))) return r;
    } else {
      c->pend_n = 0;
    }
  }
  return 0;
}

// ---------------------------------------------------------- graph replay
// Steps per replayed block when nothing is sampled (with sampling a block is
// sample_T steps, so its samples fill rows 0 .. sample_T-1 of X and its
// spectra are summed at its end).  Even: a block of ping-pong steps ends in
// the buffer set it started from.
constexpr int32_t kGraphSteps = 32;

int64_t* step_clock(const pm_ctx* c) {
  return (int64_t*)((char*)c->plan.scratch + kScratchClock);
}

// One step launched directly: fields (in place, or one ping-pong step),
// sources, samples.
int eager_step(pm_ctx* c, int64_t n, bool in_place, bool fused, bool src, bool smp, bool dft) {
  Span whole(c, 2);
  int r;
  if (fused && !in_place) {
    r = do_fused(c);
  } else {
    r = do_h(c);
    if (!r) r = do_e(c);
  }
  if (r) return r;
  const KAux& ax = c->parity ? c->aux2 : c->aux;
  if (src && smp && ax.io_merge) {
    if ((r = sample_pre(c, n, dft))) return r;
    launch_sources_sample(ax, n, n, c->stream);
    if ((r = launch_status(c, "source and sample launch"))) return r;
    return sample_post(c, n, dft);
  }
  if (src) {
    launch_sources(ax, n, c->stream);
    if ((r = launch_status(c, "source launch"))) return r;
  }
  if (smp && (r = sample_step(c, n, dft))) return r;
  return 0;
}

// Capture one block of G steps that starts in buffer set `par` into graph
// slot `slot`, on the second stream: per step the fields, the sources (table row from the step
// clock, which holds the block's first step) and the samples (row k of X);
// then the block's spectra (phasor columns from the clock) and the clock
// advanced by G.  Kernels only: nothing in the graph clears memory.
int capture_block(pm_ctx* c, int par, int slot, int32_t G, bool fused, bool src, bool smp,
                  bool dft) {
  const pm_plan& p = c->plan;
  int64_t* clk = step_clock(c);
  hipStream_t s0 = c->stream;
  const int par0 = c->parity;
  const bool span0 = c->span_on;
  c->stream = c->s2;
  c->parity = par;
  c->span_on = false;
  hipGraph_t g = nullptr;
  int r = hip_check(c, hipStreamBeginCapture(c->s2, hipStreamCaptureModeThreadLocal),
                    "begin capture");
  if (!r) {
    for (int32_t k = 0; k < G && !r; ++k) {
      if (fused) {
        r = do_fused(c);
      } else {
        r = do_h(c);
        if (!r) r = do_e(c);
      }
      if (r) break;
      const KAux& ax = c->parity ? c->aux2 : c->aux;
      if (src && smp && ax.io_merge) {
        launch_sources_sample(ax, k, k, c->stream, clk);
        r = launch_status(c, "source and sample launch");
        continue;
      }
      if (src) {
        launch_sources(ax, k, c->stream, clk);
        r = launch_status(c, "source launch");
      }
      if (!r && smp) {
        launch_sample(ax, k, c->stream);
        r = launch_status(c, "sample launch");
      }
    }
    if (!r && dft) {
      for (const KDft& d : c->dft)
        launch_dft(d, p.X, p.ncols, p.W_re, p.W_im, p.table_len, p.table_step0, G,
                   p.precision, c->stream, clk);
      r = launch_status(c, "DFT launch");
    }
    if (!r) {
      launch_clock(clk, G, true, c->stream);
      r = launch_status(c, "clock launch");
    }
    const hipError_t e = hipStreamEndCapture(c->s2, &g);
    if (!r) r = hip_check(c, e, "end capture");
  }
  c->stream = s0;
  c->parity = par0;
  c->span_on = span0;
  if (!r)
    r = hip_check(c, hipGraphInstantiate(&c->graph[slot], g, nullptr, nullptr, 0),
                  "graph instantiate");
  if (g) (void)hipGraphDestroy(g);
  return r;
}

// nrep blocks of G steps from step n0 as graph replays on the second
// stream, ordered after the caller's stream's work and before its next.
int replay_blocks(pm_ctx* c, int64_t n0, int32_t nrep, int32_t G, bool fused, bool src,
                  bool smp, bool dft) {
  // graphs per starting buffer set (fused steps) or per starting gate flag
  // set (2-D compat steps; blocks hold an even number of steps, so a block
  // ends on the set it starts with)
  const int par = fused ? c->parity : c->flag_set;
  const uint32_t key = 1u | (fused ? 2u : 0u) | (src ? 4u : 0u) | (smp ? 8u : 0u) |
                       (dft ? 16u : 0u) | ((uint32_t)G << 5);
  int r;
  if (!c->graph[par] || c->graph_key[par] != key) {
    if (c->graph[par]) {
      (void)hipGraphExecDestroy(c->graph[par]);
      c->graph[par] = nullptr;
    }
    if ((r = capture_block(c, c->parity, par, G, fused, src, smp, dft))) return r;
    c->graph_key[par] = key;
  }
  launch_clock(step_clock(c), n0, false, c->stream);
  if ((r = launch_status(c, "clock launch"))) return r;
  if ((r = hip_check(c, hipEventRecord(c->fork, c->stream), "fork record"))) return r;
  if ((r = hip_check(c, hipStreamWaitEvent(c->s2, c->fork, 0), "fork wait"))) return r;
  for (int32_t i = 0; i < nrep; ++i) {
    hipEvent_t a = c->timing ? take_event(c) : nullptr;
    if (a) (void)hipEventRecord(a, c->s2);
    if ((r = hip_check(c, hipGraphLaunch(c->graph[par], c->s2), "graph launch"))) return r;
    if (a) {
      hipEvent_t b = take_event(c);
      if (b) {
        (void)hipEventRecord(b, c->s2);
        c->spans[6].push_back({a, b});
      }
    }
  }
  if ((r = hip_check(c, hipEventRecord(c->join, c->s2), "join record"))) return r;
  return hip_check(c, hipStreamWaitEvent(c->stream, c->join, 0), "join wait");
}

// pm_run with PM_RUN_GRAPH: the odd first step of an odd fused run in place
// (as the direct path does), direct steps up to a block start, whole blocks
// replayed, the rest direct.
int run_graph(pm_ctx* c, int64_t step0, int32_t nsteps, int32_t G, bool fused, bool src,
              bool smp, bool dft) {
  int32_t s = 0;
  int r;
  auto direct = [&](int32_t k) {
    c->span_on = c->timing && (step0 + k) % c->timing_stride == 0;
    return eager_step(c, step0 + k, fused && k == 0 && (nsteps & 1), fused, src, smp, dft);
  };
  while (s < nsteps && ((fused && s == 0 && (nsteps & 1)) || (smp && (step0 + s) % G != 0))) {
    if ((r = direct(s))) return r;
    ++s;
  }
  const int32_t nrep = (nsteps - s) / G;
  if (nrep > 0) {
    // samples of a block that does not continue here are summed first (as
    // sample_step does); a block start has none of its own pending
    if (smp && c->pend_n) {
      if (dft) {
        if ((r = flush_pending(c))) return r;
      } else {
        c->pend_n = 0;
      }
    }
    if ((r = replay_blocks(c, step0 + s, nrep, G, fused, src, smp, dft))) return r;
    s += nrep * G;
  }
  while (s < nsteps) {
    if ((r = direct(s))) return r;
    ++s;
  }
  c->span_on = c->timing;
  return 0;
}

}  // namespace

int pm_run(pm_ctx* c, int64_t step0, int32_t nsteps, uint32_t flags) {
  if (!c) return PM_EINVAL;
  if (!c->configured) return fail(c, PM_ENOPLAN, "pm_run before pm_configure");
  if (nsteps < 0 || step0 < 0) return fail(c, PM_EINVAL, "negative step range");
  const pm_plan& p = c->plan;
  const bool io_only = (flags & PM_RUN_IO_ONLY) != 0;
  const bool src = (flags & PM_RUN_SOURCES) && p.n_src > 0 && !io_only;
  const bool smp = (flags & PM_RUN_SAMPLE) && p.n_sbox > 0;
  const bool dft = (flags & PM_RUN_DFT) && smp && !c->dft.empty();
  if (io_only && nsteps > 1)
    return fail(c, PM_EINVAL, "PM_RUN_IO_ONLY samples one step (nsteps %d)", nsteps);
  if (src || dft) {
    const int64_t r0 = step0 - p.table_step0;
    if (r0 < 0 || r0 + nsteps > p.table_len)
      return fail(c, PM_ERANGE, "tables cover steps [%lld, %lld), run asks [%lld, %lld)",
                  (long long)p.table_step0, (long long)(p.table_step0 + p.table_len),
                  (long long)step0, (long long)(step0 + nsteps));
  }
  if (io_only) {
    if (nsteps == 1 && smp) return sample_step(c, step0, dft);
    return 0;
  }
  int fvec = 0, fw = 0, fty = 0, fxs = 1;
  const bool fused = c->fused_ok && c->fuse && fused_shape(c, &fvec, &fw, &fty, &fxs);
  if (flags & PM_RUN_GRAPH) {
    const int32_t G = smp ? p.sample_T : kGraphSteps;
    if (G % 2) return fail(c, PM_EINVAL, "graph replay needs an even sample_T (%d)", G);
    if (!c->s2) return fail(c, PM_ENOTSUP, "graph replay needs the context's second stream");
    return run_graph(c, step0, nsteps, G, fused, src, smp, dft);
  }
  // two steps per launch for the steps that fill pairs of passes (an even
  // number of passes, so the E-side CPML memory ends in its first set); the
  // one to three steps left over go first: an odd one unfused in place, a
  // pair through the one-step kernel (an even number of ping-pong launches,
  // so the fields end in the first set too)
  const bool two = fused && c->fuse == 2 && c->fused2_ok && !smp && nsteps >= 4 &&
                   (!src || c->src_kind0);
  const int32_t head = two ? nsteps % 4 : 0;
  for (int32_t s = 0; s < nsteps;) {
    const int64_t n = step0 + s;
    c->span_on = c->timing && n % c->timing_stride == 0;
    int r;
    if (two && s >= head) {
      {
        Span whole(c, 2);
        r = do_fused2(c, n, src);
        if (r) return r;
        if (src) {
          launch_sources(c->parity ? c->aux2 : c->aux, n + 1, c->stream);
          if ((r = launch_status(c, "source launch"))) return r;
        }
      }
      s += 2;
      continue;
    }
    // fused runs keep an even number of ping-pong steps: an odd run does
    // its first step in place, so the call ends with the state in the first
    // set
    const int32_t one = two ? head : nsteps;
    if ((r = eager_step(c, n, s == 0 && (one & 1), fused, src, smp, dft))) return r;
    ++s;
  }
  c->span_on = c->timing;
  return 0;
}

int pm_dft_flush(pm_ctx* c, int64_t step_first, int32_t rows) {
  if (!c) return PM_EINVAL;
  if (!c->configured) return fail(c, PM_ENOPLAN, "pm_dft_flush before pm_configure");
  if (rows < 0 || step_first < 0) return fail(c, PM_EINVAL, "negative step or row count");
  if (rows == 0) {
    if (c->dft.empty()) {
      c->pend_n = 0;
      return 0;
    }
    return flush_pending(c);
  }
  const pm_plan& p = c->plan;
  if (c->dft.empty()) return fail(c, PM_EINVAL, "pm_dft_flush: the plan has no DFT groups");
  if (step_first % p.sample_T + rows > p.sample_T)
    return fail(c, PM_EINVAL, "rows %lld..%lld run past the %d rows of X",
                (long long)(step_first % p.sample_T),
                (long long)(step_first % p.sample_T + rows - 1), p.sample_T);
  if (c->pend_n) {
    const int64_t a0 = c->pend0, a1 = c->pend0 + c->pend_n;
    if (step_first == a0 && rows == c->pend_n) return flush_pending(c);
    if (step_first < a1 && a0 < step_first + rows)
      return fail(c, PM_EINVAL,
                  "steps %lld..%lld overlap samples pm_run has not summed (%lld..%lld)",
                  (long long)step_first, (long long)(step_first + rows - 1), (long long)a0,
                  (long long)(a1 - 1));
  }
  return flush_rows(c, step_first, rows);
}

int pm_set_row_split(pm_ctx* c, int segments) {
  if (!c) return PM_EINVAL;
  if (segments < 0 || segments > 2) return fail(c, PM_EINVAL, "segments must be 0, 1 or 2");
  if (c->s2) (void)hipStreamSynchronize(c->s2);
  drop_graphs(c);
  c->tune_xs = segments;
  return 0;
}

int pm_fused_shape(pm_ctx* c, int32_t* out) {
  if (!c || !out) return PM_EINVAL;
  out[0] = out[1] = out[2] = out[3] = 0;
  if (!c->configured || !c->fused_ok || !c->fuse) return 0;
  int vec = 0, w = 0, ty = 0, xs = 1;
  if (!fused_shape(c, &vec, &w, &ty, &xs)) return 0;
  out[0] = vec;
  out[1] = w;
  out[2] = ty;
  out[3] = xs;
  return 1;
}

int pm_row_split_active(pm_ctx* c) {
  if (!c || !c->configured || !c->fused_ok || !c->fuse) return 0;
  int vec = 0, w = 0, ty = 0, xs = 1;
  if (!fused_shape(c, &vec, &w, &ty, &xs)) return 0;
  return xs;
}

int pm_set_tuning(pm_ctx* c, int fuse, int ty, int kchunk, int vec, int order) {
  if (!c) return PM_EINVAL;
  if (ty < 2 || ty > 7 || ty == 6)
    return fail(c, PM_EINVAL, "ty must be 2, 3, 4, 5 or 7 (5 and 7: one cell per lane, 128-cell rows)");
  if (kchunk < 1) return fail(c, PM_EINVAL, "kchunk must be >= 1");
  if (vec != 0 && vec != 1 && vec != 2 && vec != 4)
    return fail(c, PM_EINVAL, "vec must be 0 (auto), 1, 2 or 4");
  if (order < 0 || order > 3) return fail(c, PM_EINVAL, "order must be 0, 1, 2 or 3");
  if (fuse < 0 || fuse > 2) return fail(c, PM_EINVAL, "fuse must be 0, 1 or 2");
  if (fuse == 2 && (ty != 4 || vec != 0))
    return fail(c, PM_EINVAL, "two steps per launch take ty 4 and vec 0 (auto)");
  // captured graphs hold the old launch shape
  if (c->s2) (void)hipStreamSynchronize(c->s2);
  drop_graphs(c);
  c->fuse = fuse;
  c->tune_ty = ty;
  c->tune_kchunk = kchunk;
  c->kchunk_set = true;
  c->tune_vec = vec;
  c->tune_order = order;
  return 0;
}

int pm_fused_active(pm_ctx* c) {
  if (!c || !c->configured) return 0;
  int vec = 0, w = 0, ty = 0;
  if (!c->fused_ok) return 0;
  int xs = 1;
  if (!(c->fuse && fused_shape(c, &vec, &w, &ty, &xs))) return 0;
  return (c->fuse == 2 && c->fused2_ok) ? 2 : 1;
}

int pm_sync(pm_ctx* c) {
  if (!c) return PM_EINVAL;
  return hip_check(c, hipStreamSynchronize(c->stream), "stream sync");
}

int pm_set_timing(pm_ctx* c, int enable) {
  if (!c) return PM_EINVAL;
  if (enable < 0) return fail(c, PM_EINVAL, "negative timing stride");
  c->timing = enable != 0;
  c->timing_stride = enable > 0 ? enable : 1;
  c->span_on = c->timing;
  c->pool_used = 0;
  for (auto& v : c->spans) v.clear();
  return 0;
}

int pm_get_timing(pm_ctx* c, int which, double* avg_ms, int64_t* count) {
  if (!c || which < 0 || which > 6 || !avg_ms || !count) return PM_EINVAL;
  int r = hip_check(c, hipStreamSynchronize(c->stream), "stream sync");
  if (!r && c->s2) r = hip_check(c, hipStreamSynchronize(c->s2), "stream sync");
  if (r) return r;
  double tot = 0;
  int64_t n = 0;
  for (auto& ab : c->spans[which]) {
    float ms = 0;
    hipError_t e = hipEventElapsedTime(&ms, ab.first, ab.second);
    if (e != hipSuccess) return hip_check(c, e, "event elapsed");
    tot += ms;
    ++n;
  }
  *avg_ms = n ? tot / n : 0.0;
  *count = n;
  return 0;
}

}  // extern "C"
