/*
 * pmbatch.h -- batched 2-D TM (Ez, Hx, Hy) time stepping of many uneven
 * grids in one launch per step, for parameter sweeps.
 *
 * The reference runs a sweep as one full Simulation per parameter set,
 * sequentially or in a process pool (src/prismo/optimization/sweep.py:
 * 119-167, ParameterSweep.run -> simulation_func), each a Python step loop
 * (src/prismo/core/simulation.py:147-164).  These entry points replace that
 * loop for a whole batch of 2-D designs at once: every design keeps its own
 * grid, materials, CPML, line sources and port sample lines; one kernel
 * advances all of them by one step; the samples of all ports of all designs
 * form one matrix whose spectra are one MFMA GEMM per block of steps.
 *
 * Scheme: the textbook Yee TM update of pmfdtd.h's PM_SCHEME_YEE in 2-D
 * (Hx -= Db dEz/dy, Hy += Db dEz/dx, Ez += Cb (dHy/dx - dHx/dy); Ez beyond
 * the grid = 0; CPML psi per derivative term; sources added after the E
 * update, samples taken after the sources).
 *
 * Conventions are those of pmfdtd.h (status codes, caller-owned device
 * memory, one stream).  Arrays of design d are x-fastest with row pitch
 * sy_d elements, ny_d rows: offset(i, j) = i + j * sy_d, zero outside each
 * component's extent (Ez (nx-1)x(ny-1), Hx (nx-1)x ny, Hy nx x (ny-1)).
 */
#ifndef PMBATCH_H
#define PMBATCH_H

#include "pmfdtd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PM_BATCH_MAX_LINES 8

/* One sampled line of a design: component (0 = Ez, 1 = Hx, 2 = Hy) at
 * x index i, rows [j0, j1), written to sample columns col .. col+(j1-j0)-1. */
typedef struct pm_batch_line {
  int32_t comp, i, j0, j1, col, pad;
} pm_batch_line;

typedef struct pm_batch_design {
  int32_t nx, ny;            /* grid nodes */
  int32_t sy, n_lines;       /* row pitch (elements, >= nx); sample lines used */
  int32_t pml[4];            /* CPML layers x-lo, x-hi, y-lo, y-hi */
  void* ez[2];               /* ping-pong field sets (state alternates every step) */
  void* hx[2];
  void* hy[2];
  void* psi_hx[2];           /* H-side CPML memory (dEz/dy at Hx, dEz/dx at Hy), ping-pong */
  void* psi_hy[2];
  void* psi_ex;              /* E-side CPML memory of Ez (dHy/dx, dHx/dy terms), in place */
  void* psi_ey;
  const void* cb;            /* Cb at Ez points: field precision, or FP16 (coeff_kind CB_F16) */
  double cb_scale;           /* FP16 Cb = cb_scale * stored value (1 otherwise) */
  double db;                 /* dt / mu0 */
  const void* prof;          /* CPML profiles, field precision: x: be ce ike bh ch ikh (nx
                                each), then y interleaved per row: 8 values per row j
                                (be ce ike bh ch ikh, two unused), ny rows */
  int32_t src_e[3];          /* Ez line source: i, j0, j1 (j0 == j1: none) */
  int32_t src_h[3];          /* Hy line source: i, j0, j1 */
  pm_batch_line line[PM_BATCH_MAX_LINES];
} pm_batch_design;

typedef struct pm_batch_plan {
  int32_t abi_version, precision;    /* PM_FP64 or PM_FP32 */
  int32_t coeff_kind, n_design;      /* PM_COEFF_CB or PM_COEFF_CB_F16 */
  double dx, dy, dt;                 /* shared by every design */
  const pm_batch_design* design;     /* host array, deep-copied */
  const double* src_e_table;         /* source values per step (device, FP64): Ez */
  const double* src_h_table;         /* Hy (may equal src_e_table scaled on the host) */
  int64_t table_step0, table_len;    /* tables cover steps [table_step0, +table_len) */
  void* X;                           /* sample ring [sample_T][ncols], field precision */
  int32_t ncols, sample_T;
  pm_dft_group dft;                  /* spectra over all sample columns */
  const void* W_re;                  /* [nfreq][table_len] phasors * dt, field precision */
  const void* W_im;
  int32_t nfreq, chunk;              /* chunk: rows per tile (>= 2) */
  void* work;                        /* device workspace, pm_batch_workspace_bytes */
  int64_t work_bytes;
} pm_batch_plan;

typedef struct pm_batch pm_batch;

int pm_batch_create(int device, void* stream, pm_batch** out);
int pm_batch_destroy(pm_batch* b);
const char* pm_batch_last_error(pm_batch* b);

/* Device workspace the plan needs (tile list, design table, step clock). */
int pm_batch_workspace_bytes(const pm_batch_plan* plan, int64_t* bytes);

/* Validate and adopt a plan; uploads the design table and tile list into
 * plan->work.  The step counter restarts at table_step0 with state in set 0. */
int pm_batch_configure(pm_batch* b, const pm_batch_plan* plan);

/* Advance every design by `nsteps` steps: H, E, sources, samples; a spectral
 * flush whenever a block of sample_T rows completes.  With PM_RUN_GRAPH the
 * whole blocks run as replays of one captured HIP graph (sample_T steps +
 * flush + clock advance; sample_T must be even), the rest eagerly. */
int pm_batch_run(pm_batch* b, int32_t nsteps, uint32_t flags);

/* Accumulate a partial block of samples (end of a run). */
int pm_batch_flush(pm_batch* b);

/* Field set (0 or 1) holding the current state; steps taken so far. */
int pm_batch_parity(pm_batch* b);
int64_t pm_batch_steps(pm_batch* b);

int pm_batch_sync(pm_batch* b);

/* Per design (n must equal the plan's n_design), waits for the stream and
 * writes bit 0 = some Ez value was Inf or NaN since pm_batch_configure.  A
 * design that diverges cannot disturb the others (their arithmetic never
 * reads it), so a caller keeps their results and reports this one failed.
 * No counterpart in the reference, which runs each design in its own
 * Simulation (src/prismo/optimization/sweep.py:190-195 catches per design). */
int pm_batch_design_status(pm_batch* b, int32_t* out, int32_t n);

/* Launch shape of the step kernel: waves (64 cells wide each) per
 * workgroup, 1, 2, 4 (default) or 8, and the workgroup order (1, default:
 * each run of 64 workgroups is grouped by XCD so neighbouring strips share
 * an L2; 0: launch order).  Rows per tile are the plan's `chunk`.  Results
 * do not depend on any of them.  Tunes what the reference's backends take as
 * fixed launch constants (src/prismo/backends/cuda_kernels.py:230-265). */
int pm_batch_set_tuning(pm_batch* b, int waves, int xcd_order);

/* Average time (ms) per step kernel (which 0) or per flush (1) since enabled. */
int pm_batch_set_timing(pm_batch* b, int enable);
int pm_batch_get_timing(pm_batch* b, int which, double* avg_ms, int64_t* count);

#ifdef __cplusplus
}
#endif
#endif /* PMBATCH_H */
