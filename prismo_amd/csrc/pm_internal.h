// Internal types shared by the libpmfdtd translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "../../include/pmfdtd.h"

namespace pm {

constexpr int kMaxAde = 8;
constexpr int kMaxSrc = 16;
constexpr int kMaxSbox = 24;
constexpr int kMaxDft = 8;

// Fixed block shape of the stencil kernels: one wave64 per x-row, 4 rows.
constexpr int kBX = 64;
constexpr int kBY = 4;

// Dispersive box as the kernels use it: the time-centred coupled update
// derived at configure from the reference's coefficients (see ade_cell).
struct KAde {
  int comp, kind, i0, i1, j0, j1, k0, k1;
  double ke, kp, kq, ks;      // E+ = ke E + kp P - kq Q + ks curlH
  double a0, a1, a2, a3, a4;  // P+ = a0 E+ + a1 E + a2 P + Q;  Q+ = a3 P + a4 E
  void* p;                    // P (Lorentz) or J (Drude), box order
  void* q;                    // Q (Lorentz only)
};

struct KSrc {
  int comp, kind, i0, i1, j0, j1, k0, k1;
  const double* table;
  const double* profile;
  double divisor;
};

struct KSbox {
  int comp, i0, i1, j0, j1, k0, k1, col0;
};

struct KDft {
  int col0, col1, f0, f1;
  void* acc_re;
  void* acc_im;
  void* cmp_re;
  void* cmp_im;
};

// Everything a stencil kernel needs, passed by value.
struct KGrid {
  int nx, ny, nz, ndim;
  int64_t sy, sz;
  double dx, dy, dz, dt;
  double idx, idy, idz;
  void* E[3];
  void* H[3];
  const void* ca[3];
  const void* cb[3];
  const void* da[3];
  const void* db[3];
  double cb_scale, db_const;
  int pml[6];
  const void* pml_be[3];
  const void* pml_ce[3];
  const void* pml_ike[3];
  const void* pml_bh[3];
  const void* pml_ch[3];
  const void* pml_ikh[3];
  void* psi_e[3][2];
  void* psi_h[3][2];
  int has_pml;
  int n_ade;
  int ade_j0, ade_j1, ade_k0, ade_k1;  // union of the boxes' j and k ranges
  const KAde* ade;  // device copy of the ADE boxes (in scratch)
  int* flags;       // compat gating flags of this step (scratch)
  int* flags_next;  // the other set, cleared by the 2-D compat H kernel
};

// scratch layout (caller-owned, PM_SCRATCH_BYTES)
constexpr int kScratchFlags = 0;
constexpr int kScratchClock = 512;  // step clock of graph-replayed runs (int64)
constexpr int kScratchAde = 1024;
constexpr int kScratchSbox = 2048;  // sample boxes (merged sampling launch)
constexpr int kScratchSrc = 2816;   // source boxes (merged source launch)

// One dispersive cell: new E from the old E and the curl of H, advancing the
// polarisation state at linear index l of the box.  Drude boxes have kq = 0,
// a3 = a4 = 0 and no Q array.
// Global-memory pointer: pointers read from device tables are generic to
// the compiler, which then emits flat accesses whose completion it cannot
// count per access (they may be LDS); every such table here points to
// device memory.
template <typename T>
using gptr = T __attribute__((address_space(1)))*;

// The same with the state (p0 = P[l], q0 = Q[l]; q0 unused for Drude)
// already loaded by the caller, which can then issue those loads early.
template <typename T>
__device__ __forceinline__ T ade_cell_state(const KAde& a, int64_t l, T p0, T q0, T curl, T e) {
  const gptr<T> P = (gptr<T>)a.p;
  if (a.kind == 1) {
    const gptr<T> Q = (gptr<T>)a.q;
    const T en = (T)a.ke * e + (T)a.kp * p0 - (T)a.kq * q0 + (T)a.ks * curl;
    P[l] = (T)a.a0 * en + (T)a.a1 * e + (T)a.a2 * p0 + q0;
    Q[l] = (T)a.a3 * p0 + (T)a.a4 * e;
    return en;
  }
  const T en = (T)a.ke * e + (T)a.kp * p0 + (T)a.ks * curl;
  P[l] = (T)a.a0 * en + (T)a.a1 * e + (T)a.a2 * p0;
  return en;
}

// The same with the nine coefficients already in the field type, cf = ke
// kp kq ks a0 a1 a2 a3 a4 (the values (T)a.ke ... (T)a.a4).
template <typename T>
__device__ __forceinline__ T ade_cell_coef(const KAde& a, const T* cf, int64_t l, T p0, T q0,
                                           T curl, T e) {
  const gptr<T> P = (gptr<T>)a.p;
  if (a.kind == 1) {
    const gptr<T> Q = (gptr<T>)a.q;
    const T en = cf[0] * e + cf[1] * p0 - cf[2] * q0 + cf[3] * curl;
    P[l] = cf[4] * en + cf[5] * e + cf[6] * p0 + q0;
    Q[l] = cf[7] * p0 + cf[8] * e;
    return en;
  }
  const T en = cf[0] * e + cf[1] * p0 + cf[3] * curl;
  P[l] = cf[4] * en + cf[5] * e + cf[6] * p0;
  return en;
}

template <typename T>
__device__ __forceinline__ T ade_cell(const KAde& a, int64_t l, T curl, T e) {
  const T p0 = ((gptr<const T>)a.p)[l];
  const T q0 = a.kind == 1 ? ((gptr<const T>)a.q)[l] : (T)0;
  return ade_cell_state<T>(a, l, p0, q0, curl, e);
}

struct KAux {
  int nx, ny, nz;
  int64_t sy, sz;
  void* F[6];      // Ex Ey Ez Hx Hy Hz
  int precision;
  int n_src;
  KSrc src[kMaxSrc];
  int64_t table_step0, table_len;
  int n_sbox, sample_T, ncols;
  KSbox sbox[kMaxSbox];
  void* X;
  // device copies of the box tables (caller scratch) for one launch over
  // all boxes of a kind; src_merge is false when two source boxes of one
  // component overlap (their adds then keep the reference's order, one
  // launch per box)
  const KSbox* sbox_dev;
  const KSrc* src_dev;
  int64_t sbox_max, src_max;  // cells of the largest box
  bool src_merge;
  // sources and samples in one launch: merged sources, and no sample box
  // meets a source box of its component (a sample then never reads a cell
  // the same launch adds to)
  bool io_merge;
};

// One step's buffer pair for the fused kernel.
struct FusedBufs {
  const void* Ei[3];
  const void* Hi[3];
  void* Eo[3];
  void* Ho[3];
  const void* psh_in[3][2];
  void* psh_out[3][2];
  // two-step pass only: E-side CPML memory written (read: KGrid::psi_e)
  void* pse_out[3][2];
};

// Launchers (defined in the .hip files).
void launch_yee_h(const KGrid& g, int precision, int coeff_kind, hipStream_t s);
void launch_yee_e(const KGrid& g, int precision, int coeff_kind, hipStream_t s);
void launch_compat_flags2d(const KGrid& g, hipStream_t s);
void launch_compat_flags_clear(int* flags, hipStream_t s);
void launch_compat_h(const KGrid& g, int coeff_kind, hipStream_t s);
void launch_compat_e(const KGrid& g, int coeff_kind, hipStream_t s);
// One fused 2-D TM step (tm2d.hip): Ez, Hx, Hy of set "i" -> set "o", the
// H-side CPML memory of the x (Hy) and y (Hx) slabs likewise, the E-side
// memory of the plan updated in place; rows per tile `chunk`.
struct Tm2dBufs {
  const void* Ei;
  const void* Hi[2];
  void* Eo;
  void* Ho[2];
  const void* psh_in[2];
  void* psh_out[2];
};
int launch_tm2d_fused(const KGrid& g, int precision, int ck, const Tm2dBufs& b, int chunk,
                      hipStream_t s);
// xs: row segments per padded row (1, or 2: rows of 2 x 64 vec w cells).
int launch_yee_fused_f32(const KGrid& g, int ck, const FusedBufs& fb, int vec, int w, int ty,
                         int kchunk, int order, int xs, hipStream_t s);
int launch_yee_fused_f64(const KGrid& g, int ck, const FusedBufs& fb, int vec, int w, int ty,
                         int kchunk, int order, int xs, hipStream_t s);
bool fused_supported(int precision, int ck, int vec, int w, int ty, int xs);
// Two Yee steps per launch (fused2_kernel.h), FP32 fields, Cb in FP32 or
// FP16, no dispersive boxes; the sources of the first step (table row
// src_row, n_src entries of the device table src) are added inside.
// The slab tiles go to stream s2 (forked from s with event fork, joined back
// with event join) so they run beside the interior tiles; s2 = NULL keeps
// both launches on s.
int launch_yee_fused2_f32(const KGrid& g, int ck, const FusedBufs& fb, int vec, int w, int ty,
                          int kchunk, int order, int64_t src_row, int n_src, const KSrc* src,
                          hipStream_t s, hipStream_t s2, hipEvent_t fork, hipEvent_t join);
bool fused2_supported(int vec, int w, int ty);
bool fused2_grid_ok(const KGrid& g);  // x CPML layers fit the kernel's LDS slots
// The plan's arrays are laid out as the fused kernel addresses them
// (packed field sets, Cb, CPML memory pairs and profiles within 4 GiB).
bool fused_addressable(const KGrid& g, const FusedBufs& fb, int esz);
// With `clock`, `step` is relative: the table row is step + *clock - table_step0.
void launch_sources(const KAux& a, int64_t step, hipStream_t s,
                    const int64_t* clock = nullptr);
// *clock = v (add false) or *clock += v, in one thread of a kernel (so it can
// be captured in a graph).
void launch_clock(int64_t* clock, int64_t v, bool add, hipStream_t s);
void launch_sample(const KAux& a, int64_t step, hipStream_t s);
// Sources of step `step` (relative with `clock`, as launch_sources) and the
// samples into X row `row`, one launch (a.io_merge must hold).
void launch_sources_sample(const KAux& a, int64_t step, int64_t row, hipStream_t s,
                           const int64_t* clock = nullptr);
void launch_dft(const KDft& d, const void* X, int ncols, const void* W_re, const void* W_im,
                int64_t table_len, int64_t wcol0, int rows, int precision, hipStream_t s,
                const int64_t* wclock = nullptr);

}  // namespace pm
