/*
 * pmfdtd.h -- C ABI of libpmfdtd, the MI355X (gfx950) FDTD time-step library.
 *
 * The reference (rithulkamesh/prismo) is pure Python: its time step is
 * MaxwellUpdater.step (src/prismo/core/solver.py:535-552) driven by
 * FDTDSolver.run_steps (solver.py:643-662) and Simulation.step
 * (src/prismo/core/simulation.py:147-164), with sources
 * (src/prismo/sources/base.py:141-157) and monitors
 * (src/prismo/monitors/base.py:136-150) called from Python every step.  It has
 * no FFI; every entry point below is the native body that a 'mi355x' backend
 * binds in place of those Python calls (through ctypes: prismo_amd/_native.py).
 *
 * Conventions
 *  - Status: 0 = ok, > 0 = a hipError_t, < 0 = a PM_E* argument/config error.
 *    pm_last_error() gives the message of the last failure on that context.
 *  - Memory: every device buffer is owned by the caller (torch) and borrowed
 *    from pm_configure until the next pm_configure or pm_destroy.  The library
 *    allocates nothing on the device.
 *  - Streams: all work is enqueued on the stream given to pm_create (torch's
 *    current stream); calls return before the GPU finishes.  One context per
 *    host thread.
 *  - Layout: every field-shaped buffer is x-fastest,
 *      offset(i, j, k) = i + j * sy + k * sz      (elements),
 *    with one pitch (sy, sz) shared by all six components, and is zero outside
 *    the component's reference extent (src/prismo/core/grid.py:137-176).
 */
#ifndef PMFDTD_H
#define PMFDTD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PM_ABI_VERSION 7
#define PM_SCRATCH_BYTES 4096

/* error codes (< 0) */
#define PM_EINVAL   (-1)  /* bad argument / inconsistent plan */
#define PM_ENOTSUP  (-2)  /* combination not built into this library */
#define PM_ENOPLAN  (-3)  /* pm_run before pm_configure */
#define PM_ERANGE   (-4)  /* step index outside the uploaded tables */

enum pm_component { PM_EX = 0, PM_EY, PM_EZ, PM_HX, PM_HY, PM_HZ };

enum pm_scheme {
  /* Reference arithmetic of src/prismo/core/solver.py:167-533, operation for
   * operation (forward H differences, trimmed ranges, 2-D gating, per-step
   * coefficient averaging evaluated in the same order).  FP64 only. */
  PM_SCHEME_COMPAT = 0,
  /* Textbook Yee leapfrog on the same staggering (backward H differences,
   * standard 2-D TM signs) with precomputed coefficients, CPML and ADE. */
  PM_SCHEME_YEE = 1
};

enum pm_precision { PM_FP64 = 0, PM_FP32 = 1 };

enum pm_coeff_kind {
  /* per-component coefficient arrays ca/cb/da/db in the field precision */
  PM_COEFF_FULL = 0,
  /* lossless non-magnetic media (sigma_e = sigma_m = 0, mu_r = 1, so
   * Ca = Da = 1 exactly and Db = db_const): only cb[3] in field precision */
  PM_COEFF_CB = 1,
  /* as PM_COEFF_CB with cb[3] stored as FP16 holding Cb / cb_scale;
   * FP32 fields only */
  PM_COEFF_CB_F16 = 2
};

/* Additive source term: field[c] += value for every cell c of the box, after
 * the E update of each step (src/prismo/core/simulation.py:158-160).
 *   kind 0: value = table[n]                                  (point, plane)
 *   kind 1: value = (table[n] * profile[c]) / divisor  (divisor 0 = none)
 * profile is in box order (i fastest).  table is indexed by the absolute
 * step index n minus pm_plan.table_step0. */
typedef struct pm_source {
  int32_t comp, kind;
  int32_t i0, i1, j0, j1, k0, k1;
  const double* table;
  const double* profile;
  double divisor;
} pm_source;

/* One monitored component box, copied each step into row (n mod T) of the
 * sample matrix X[T][ncols] (field precision) starting at column col0, box
 * order i fastest.  Spectra: pm_dft_group. */
typedef struct pm_sample_box {
  int32_t comp;
  int32_t i0, i1, j0, j1, k0, k1;
  int32_t col0;
} pm_sample_box;

/* Spectral accumulation over sample columns [col0, col1) and phasor rows
 * [f0, f1): F[f][p] += sum_n W[f][n] * X[n][p], complex W, real X.
 * acc_re/acc_im are [(f1-f0)][(col1-col0)] FP64 (plain) or FP32 with
 * cmp_re/cmp_im compensation terms (precision form). */
typedef struct pm_dft_group {
  int32_t col0, col1, f0, f1;
  void* acc_re;
  void* acc_im;
  void* cmp_re;
  void* cmp_im;
} pm_dft_group;

/* Dispersive (ADE) region for one E component: a box of cells with its own
 * state arrays in box order (zero-filled by the caller before the first
 * step).  c0..c3 are the coefficients exactly as
 * src/prismo/materials/dispersion.py:189-231 (Lorentz, one pole) and
 * 267-286 (Drude) return them; P and J are in units of eps0.  The library
 * couples them to the E update with the drive averaged over the half steps
 * on either side (time-centred), so the discrete eps(w) matches the closed
 * form to O((w dt)^2):
 *   Lorentz: P+ = c2 P + c3 P- + [(c0 E+ + c1 E) + (c0 E + c1 E-)] / 2
 *   Drude:   J+ = c1 J + c0 (E+ + E) / 2
 *   eps_inf (E+ - E) + (P+ - P) = cb_scale curlH,  P+ - P = dt (J+ + J) / 2. */
typedef struct pm_ade_box {
  int32_t comp, kind;            /* comp in 0..2; kind 1 Lorentz, 2 Drude */
  int32_t i0, i1, j0, j1, k0, k1;
  double eps_inf;                /* > 0 */
  double c0, c1, c2, c3;          /* Lorentz C0..C3; Drude C0, C1 (c2, c3 unused) */
  double cb_scale;                /* dt / eps0 (curl H -> E units), > 0 */
  void* p;                        /* P (Lorentz) or J (Drude) */
  void* p_prev;                   /* Lorentz: second state, Q = c3 P- + c1 E- / 2 */
} pm_ade_box;

typedef struct pm_plan {
  int32_t abi_version;            /* = PM_ABI_VERSION */
  int32_t ndim;                   /* 2 (TM: Ez, Hx, Hy in yee) or 3 */
  int32_t scheme, precision, coeff_kind, pad0;
  int32_t nx, ny, nz;             /* total node counts (grid.dimensions) */
  int32_t pad1;
  int64_t sy, sz;                 /* element strides of the padded layout */
  double dx, dy, dz, dt;

  void* E[3];
  void* H[3];
  /* Coefficients at each component's own sample points, field layout:
   * ca/cb for Ex, Ey, Ez and da/db for Hx, Hy, Hz.  Both schemes take the
   * averages of the cell-centred Ca, Cb, Da, Db of solver.py:113-133 that
   * solver.py:458-533 forms (host-computed once, in the same summation order,
   * so compat stays bit-identical).  Unused entries may be NULL. */
  const void* ca[3];
  const void* cb[3];
  const void* da[3];
  const void* db[3];
  double cb_scale;                /* PM_COEFF_CB_F16: Cb = cb_scale * cb[c] */
  double db_const;                /* PM_COEFF_CB_F16: uniform Db */

  /* CPML (yee only).  Thickness per face x-,x+,y-,y+,z-,z+ in cells.
   * Profiles per axis a (length n_a): *_e at half nodes (E-update terms),
   * *_h at nodes (H-update terms): b = recursion decay, c = recursion gain,
   * ik = 1/kappa.  psi buffers: psi_e[a][q] / psi_h[a][q], q = 0,1 the two
   * field components whose update differentiates along axis a; each is a
   * slab array of extent 2*t_a along a (low slab then high slab) by the full
   * extent of the other two axes, x-fastest. */
  int32_t pml[6];
  const void* pml_be[3];
  const void* pml_ce[3];
  const void* pml_ike[3];
  const void* pml_bh[3];
  const void* pml_ch[3];
  const void* pml_ikh[3];
  void* psi_e[3][2];
  void* psi_h[3][2];

  /* dispersive regions */
  int32_t n_ade, pad2;
  const pm_ade_box* ade;          /* host array, copied at configure */

  /* sources */
  int32_t n_src, pad3;
  const pm_source* src;           /* host array, copied at configure */
  int64_t table_step0;            /* absolute step of table[0] */
  int64_t table_len;

  /* monitors */
  int32_t n_sbox, sample_T;       /* sample_T = rows of X (flush period) */
  const pm_sample_box* sbox;      /* host array */
  int32_t ncols, n_dft;
  void* X;                        /* [sample_T][ncols] field precision */
  const pm_dft_group* dft;        /* host array */
  const void* W_re;               /* [nfreq_total][table_len] phasors * dt */
  const void* W_im;               /*  (FP64 plain, FP32 precision) */
  int32_t nfreq_total, pad4;

  /* scratch owned by the caller: >= PM_SCRATCH_BYTES, 256-byte aligned
   * (gating flags and the device copy of the ADE box table) */
  void* scratch;

  /* Optional second buffer set for the fused step (yee; 3-D with padded
   * row sy in {128, 256, 512}, or 2-D TM without dispersive boxes or full
   * coefficient arrays).  With E2/H2 and
   * psi_h2 (for every psi_h given) present, pm_run advances steps with one
   * fused H+E kernel that reads one set and writes the other; a run with an
   * odd step count does its first step in place, so every call returns with
   * the state in E/H/psi_h.  Same layout and extents as E/H/psi_h. */
  void* E2[3];
  void* H2[3];
  void* psi_h2[3][2];
  /* Optional: a second E-side CPML memory set (same layout as psi_e).  With
   * it (for every psi_e given), FP32 plans with Cb in FP32 or FP16, no
   * dispersive boxes, padded rows of 128 or 256 cells and at most 64 x CPML
   * layers can advance two steps per launch (pm_set_tuning fuse = 2). */
  void* psi_e2[3][2];
} pm_plan;

typedef struct pm_ctx pm_ctx;

/* Create a context on HIP device `device` enqueuing on `stream`
 * (hipStream_t, NULL = the legacy default stream). */
int pm_create(int device, void* stream, pm_ctx** out);
int pm_destroy(pm_ctx* ctx);
const char* pm_last_error(pm_ctx* ctx);
int pm_abi_version(void);

/* Validate and adopt a plan (deep-copies the host arrays it points to). */
int pm_configure(pm_ctx* ctx, const pm_plan* plan);

/* pm_configure for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  plan->scratch must be NULL: the library
 * uses its own PM_SCRATCH_BYTES (4 KiB) on the context's device instead (at
 * most 8 dispersive boxes, 24 monitor boxes and 16 sources, as for
 * pm_configure; more is refused with PM_EINVAL).  Allocates no device memory.
 * That scratch is one array per device, shared by every context configured
 * this way on it and keeping the last configure's tables: such contexts on
 * one device must not be configured or run concurrently (they race), and a
 * context must be configured again after another one was. */
int pm_configure_static(pm_ctx* ctx, const pm_plan* plan);

/* One H half-step (MaxwellUpdater.update_magnetic_fields,
 * solver.py:135-149) / one E half-step (update_electric_fields, 151-165).
 * No sources or monitors. */
int pm_update_h(pm_ctx* ctx);
int pm_update_e(pm_ctx* ctx);

/* `nsteps` full steps starting at absolute step index step0 (the step that
 * produces time t_{step0+1}).  Per step: H update, E update, sources with
 * table row step0+s (flag PM_RUN_SOURCES), sampling into X row
 * (step0+s) mod sample_T (flag PM_RUN_SAMPLE), and a spectral flush of the
 * completed block whenever (step0+s+1) mod sample_T == 0 (flag PM_RUN_DFT). */
#define PM_RUN_SOURCES 1u
#define PM_RUN_SAMPLE  2u
#define PM_RUN_DFT     4u
/* PM_RUN_GRAPH: whole blocks of G steps (G = sample_T with PM_RUN_SAMPLE,
 * else 32; with sampling a block starts at a step that is a multiple of
 * sample_T) run as replays of a HIP graph captured on the context's second
 * stream the first time (one per starting buffer set and flag combination,
 * kept until the next pm_configure / pm_set_tuning), ordered after the
 * context stream's work and before its next; the other steps of the run are
 * launched directly.  Each replay reads its first step from a device step
 * clock in the plan's scratch (offset 512) for the source and phasor tables.
 * The results are bitwise those of the same run without the flag.  Two steps
 * per launch (pm_set_tuning fuse 2) are not used in a graph.  PM_EINVAL when
 * G is odd.  Replaces the reference's Python step loop
 * (src/prismo/core/simulation.py:147-164). */
#define PM_RUN_GRAPH   8u
/* Take no step: sample step step0 (nsteps must be 1) from the fields as
 * they are, with PM_RUN_SAMPLE / PM_RUN_DFT as above.  For callers that add
 * terms of their own after pm_run's step and sources (src/prismo/core/
 * simulation.py:158-164: every source before any monitor). */
#define PM_RUN_IO_ONLY 16u
int pm_run(pm_ctx* ctx, int64_t step0, int32_t nsteps, uint32_t flags);

/* Spectral flush outside pm_run.
 *  rows == 0: sum the samples pm_run has taken but not yet summed (the
 *    partial block a run leaves when it ends mid-block); step_first is not
 *    used.
 *  rows > 0: sum rows (step_first mod sample_T) .. + rows - 1 of X as the
 *    samples of steps [step_first, step_first + rows) (phasor columns
 *    step_first - table_step0 ..), e.g. samples the caller wrote into X
 *    itself.  The rows must not run past the end of X (PM_EINVAL) and the
 *    phasor tables must cover the steps (PM_ERANGE); a range equal to pm_run's
 *    pending samples sums and clears them, one that only overlaps them is
 *    refused (PM_EINVAL). */
int pm_dft_flush(pm_ctx* ctx, int64_t step_first, int32_t rows);

/* Launch shape of the fused step: rows per workgroup (ty in {2, 3, 4, 5, 7}),
 * z planes per workgroup (kchunk >= 1; 2-D plans: rows per tile, default 3,
 * ty, vec and order unused), cells per lane (vec in {1, 2, 4}, 0 = auto)
 * and tile order (0: y-tiles fastest, 1: z-chunks fastest, each XCD taking a
 * contiguous range; 2: y-tiles fastest in hardware round-robin; 3: as 2 in
 * groups of 64 tiles of which each XCD takes 8 neighbours); fuse = 0 forces
 * the unfused kernels, fuse = 2 asks for two steps per launch (temporal
 * blocking; ty must be 4 and vec 0; runs that sample monitors or add
 * sources with a per-cell profile, and the one to three steps of a run that
 * do not fill pairs of passes, take the one-step kernel).  Defaults: 1, 4,
 * 8, 0, 2.  Used by the autotuner. */
int pm_set_tuning(pm_ctx* ctx, int fuse, int ty, int kchunk, int vec, int order);

/* Rows of the fused one-step kernel in workgroups of a whole padded row
 * (segments = 1) or of half of it (2: the E update of a segment's last cell
 * recomputes H of the cell to its right; plans with Cb-only coefficients
 * whose middle cell lies outside the x CPML layers); 0 (default) = whole
 * rows when a launch shape fits them, else two segments (FP64 512-cell
 * rows).  The results do not depend on it.  pm_row_split_active: segments
 * the next fused step takes (0 when it is not fused).  Used by the
 * autotuner; no counterpart in the reference. */
int pm_set_row_split(pm_ctx* ctx, int segments);
int pm_row_split_active(pm_ctx* ctx);
/* Launch shape the next one-step fused launch takes: out[0..3] = cells per
 * lane, waves per row segment, rows per workgroup, row segments (the
 * tuning asked for, or what it falls back to); returns 1, or 0 (out all 0)
 * when the step is not fused. */
int pm_fused_shape(pm_ctx* ctx, int32_t* out);

/* Steps per fused launch pm_run will use with the current plan and tuning
 * for a run without monitor sampling: 0 (unfused kernels), 1 or 2. */
int pm_fused_active(pm_ctx* ctx);

/* Block until all work on the context's stream has finished. */
int pm_sync(pm_ctx* ctx);

/* Average kernel time (ms) since timing was enabled, measured with HIP
 * events on the context stream around every launch of kernel class `which`
 * (0 = H update, 1 = E update, 2 = whole step, 3 = fused H+E step,
 * 4 = spectral flush of one sample block, 5 = two-step launch, 6 = one graph
 * replay of a PM_RUN_GRAPH block, timed on the second stream); enabled (and reset) by
 * pm_set_timing.  enable = 0 turns timing off; enable = s >= 1 times the
 * kernels of every s-th step of pm_run (steps whose index is a multiple of
 * s), so a long run can be timed with few events. */
int pm_set_timing(pm_ctx* ctx, int enable);
int pm_get_timing(pm_ctx* ctx, int which, double* avg_ms, int64_t* count);

#ifdef __cplusplus
}
#endif
#endif /* PMFDTD_H */
