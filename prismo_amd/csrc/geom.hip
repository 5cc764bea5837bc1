// Geometry -> coefficient build on the device (include/pmgeom.h).
//
// k_raster: one lane per grid node, x fastest; every shape is tested in list
// order and the last one holding the node sets its index (painter's order).
// The tests are the reference's FP64 expressions in its operation order
// (src/prismo/geometry/shapes.py:122-128 box, 153-157 sphere, 199-212
// cylinder, 243-282 polygon ray casting); this unit is built with
// -ffp-contract=off so no multiply-add is fused, and sqrt / division are the
// correctly rounded IEEE operations numpy uses, so the result equals the
// reference's boolean masks exactly.
//
// k_cb: one lane per sample point of one E component; the four (2-D Ex/Ey:
// two) node values come from the per-material table and are averaged in the
// reference's summation order (src/prismo/core/solver.py:458-533), then
// stored as FP64, FP32 (rounded to nearest) or FP16 of Cb / cb_scale.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/pmfdtd.h"
#include "../../include/pmgeom.h"

namespace pm {
namespace {

__device__ __forceinline__ bool in_box(const pm_shape& s, double px, double py, double pz) {
  return fabs(px - s.center[0]) <= s.half[0] && fabs(py - s.center[1]) <= s.half[1] &&
         fabs(pz - s.center[2]) <= s.half[2];
}

__device__ __forceinline__ bool in_sphere(const pm_shape& s, double px, double py, double pz) {
  const double dx = px - s.center[0], dy = py - s.center[1], dz = pz - s.center[2];
  // np.linalg.norm over a row of three: sqrt((dx^2 + dy^2) + dz^2)
  const double ss = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
  return __dsqrt_rn(ss) <= s.radius;
}

__device__ __forceinline__ bool in_cylinder(const pm_shape& s, double px, double py,
                                            double pz) {
  const double p[3] = {px, py, pz};
  const int a = s.axis;
  const int u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
  const double axial = fabs(p[a] - s.center[a]);
  const double du = p[u] - s.center[u], dv = p[v] - s.center[v];
  const double radial = __dsqrt_rn(__dadd_rn(__dmul_rn(du, du), __dmul_rn(dv, dv)));
  return radial <= s.radius && axial <= s.half_height;
}

__device__ bool in_polygon(const pm_shape& s, const double* __restrict__ verts, double px,
                           double py, double pz) {
  if (!(pz >= s.zmin && pz <= s.zmax)) return false;
  const double* v = verts + 2 * s.vert0;
  const int n = s.n_vert;
  int count = 0;
  for (int e = 0; e < n; ++e) {
    const int f = e + 1 == n ? 0 : e + 1;
    const double x1 = v[2 * e], y1 = v[2 * e + 1];
    const double x2 = v[2 * f], y2 = v[2 * f + 1];
    if ((y1 > py) != (y2 > py)) {
      // xi < (v2x - v1x) * (yi - v1y) / (v2y - v1y) + v1x, left to right
      const double xc = __dadd_rn(__ddiv_rn(__dmul_rn(x2 - x1, py - y1), y2 - y1), x1);
      if (px < xc) ++count;
    }
  }
  return (count & 1) != 0;
}

__global__ __launch_bounds__(256) void k_raster(int nx, int ny, int nz,
                                                const double* __restrict__ x,
                                                const double* __restrict__ y,
                                                const double* __restrict__ z,
                                                const pm_shape* __restrict__ shapes,
                                                int n_shapes, const double* __restrict__ verts,
                                                uint8_t* __restrict__ out) {
  const int64_t n = (int64_t)nx * ny * nz;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(c % nx);
    const int64_t r = c / nx;
    const int j = (int)(r % ny), k = (int)(r / ny);
    const double px = x[i], py = y[j], pz = z[k];
    int m = 0;
    for (int t = 0; t < n_shapes; ++t) {
      const pm_shape& s = shapes[t];
      bool in = false;
      switch (s.kind) {
        case PM_SHAPE_BOX: in = in_box(s, px, py, pz); break;
        case PM_SHAPE_SPHERE: in = in_sphere(s, px, py, pz); break;
        case PM_SHAPE_CYLINDER: in = in_cylinder(s, px, py, pz); break;
        case PM_SHAPE_POLYGON: in = in_polygon(s, verts, px, py, pz); break;
        default: break;
      }
      if (in) m = s.material;
    }
    out[c] = (uint8_t)m;
  }
}

// FP64 -> FP16 bits, round to nearest even, directly from the double (one
// rounding, as numpy converts)
__device__ __forceinline__ uint16_t d2h_rne(double d) {
  const uint64_t b = (uint64_t)__double_as_longlong(d);
  const uint16_t sg = (uint16_t)((b >> 48) & 0x8000u);
  const uint64_t a = b & 0x7fffffffffffffffull;
  if (a >= 0x7ff0000000000000ull) return sg | (a == 0x7ff0000000000000ull ? 0x7c00 : 0x7e00);
  const int e = (int)(a >> 52) - 1023;
  if (e >= 16) return sg | 0x7c00;
  if (e < -25) return sg;
  const uint64_t m = (a & 0xfffffffffffffull) | 0x10000000000000ull;  // 53 bits
  const int shift = e >= -14 ? 42 : 42 + (-14 - e);                     // 42 .. 53
  uint64_t q = m >> shift;
  const uint64_t rem = m & ((1ull << shift) - 1ull), half = 1ull << (shift - 1);
  if (rem > half || (rem == half && (q & 1ull))) ++q;
  if (e >= -14) return sg | (uint16_t)(((uint64_t)(e + 15) << 10) + (q - 1024ull));
  return sg | (uint16_t)q;  // subnormal (a carry into 1024 is the smallest normal)
}

__global__ __launch_bounds__(256) void k_cb(int nx, int ny, int nz, int comp,
                                            const uint8_t* __restrict__ idx,
                                            const double* __restrict__ lut, int n_lut,
                                            int storage, double cb_scale, int64_t sy,
                                            int64_t sz, void* __restrict__ out,
                                            int* __restrict__ bad) {
  const bool two_d = nz == 1;
  // component extents (grid.py get_field_shape): the axis it points along
  // keeps all nodes, each other axis loses one (2-D: z has one plane)
  const int cx = comp == 0 ? nx : nx - 1;
  const int cy = comp == 1 ? ny : ny - 1;
  const int cz = two_d ? 1 : (comp == 2 ? nz : nz - 1);
  const int64_t n = (int64_t)cx * cy * cz;
  auto node = [&](int i, int j, int k) -> double {
    int m = idx[(int64_t)i + (int64_t)nx * ((int64_t)j + (int64_t)ny * k)];
    if (m >= n_lut) {
      if (bad) *bad = 1;
      m = 0;
    }
    return lut[m];
  };
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(c % cx);
    const int64_t r = c / cx;
    const int j = (int)(r % cy), k = (int)(r / cy);
    double v;
    if (two_d && comp != 2) {
      // Ex: 0.5 * (A[:, :-1] + A[:, 1:]); Ey: 0.5 * (A[:-1, :] + A[1:, :])
      const double a = node(i, j, 0);
      const double b = comp == 0 ? node(i, j + 1, 0) : node(i + 1, j, 0);
      v = __dmul_rn(0.5, __dadd_rn(a, b));
    } else {
      double a, b, e, f;
      if (comp == 0) {  // A[:, :-1, :-1] + A[:, 1:, :-1] + A[:, :-1, 1:] + A[:, 1:, 1:]
        a = node(i, j, k); b = node(i, j + 1, k); e = node(i, j, k + 1); f = node(i, j + 1, k + 1);
      } else if (comp == 1) {  // A[:-1, :, :-1] + A[1:, :, :-1] + A[:-1, :, 1:] + A[1:, :, 1:]
        a = node(i, j, k); b = node(i + 1, j, k); e = node(i, j, k + 1); f = node(i + 1, j, k + 1);
      } else {  // A[:-1, :-1] + A[1:, :-1] + A[:-1, 1:] + A[1:, 1:]
        a = node(i, j, k); b = node(i + 1, j, k); e = node(i, j + 1, k); f = node(i + 1, j + 1, k);
      }
      v = __dmul_rn(0.25, __dadd_rn(__dadd_rn(__dadd_rn(a, b), e), f));
    }
    const int64_t o = (int64_t)i + (int64_t)j * sy + (int64_t)k * sz;
    if (storage == PM_CB_F64)
      ((double*)out)[o] = v;
    else if (storage == PM_CB_F32)
      ((float*)out)[o] = (float)v;
    else
      ((uint16_t*)out)[o] = d2h_rne(__ddiv_rn(v, cb_scale));
  }
}

int blocks(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return (int)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

int status() {
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

}  // namespace
}  // namespace pm

using namespace pm;

extern "C" {

int pm_raster(void* stream, int32_t nx, int32_t ny, int32_t nz, const double* x,
              const double* y, const double* z, const pm_shape* shapes, int32_t n_shapes,
              const double* verts, int64_t n_verts, uint8_t* out) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || n_shapes < 0 || !x || !y || !z || !out) return PM_EINVAL;
  if (n_shapes > 0 && !shapes) return PM_EINVAL;
  (void)n_verts;  // vertex ranges are checked by the caller that packs them
  const int64_t n = (int64_t)nx * ny * nz;
  hipLaunchKernelGGL(k_raster, dim3(blocks(n)), dim3(256), 0, (hipStream_t)stream, nx, ny, nz,
                     x, y, z, shapes, n_shapes, verts, out);
  return status();
}

int pm_cb_from_index(void* stream, int32_t nx, int32_t ny, int32_t nz, const uint8_t* index,
                     const double* cb_lut, int32_t n_lut, int32_t comp, int32_t storage,
                     double cb_scale, int64_t sy, int64_t sz, void* out, int32_t* bad_index) {
  if (nx < 2 || ny < 2 || nz < 1 || !index || !cb_lut || n_lut < 1 || n_lut > 256 || !out)
    return PM_EINVAL;
  if (comp < 0 || comp > 2 || storage < PM_CB_F64 || storage > PM_CB_F16) return PM_EINVAL;
  if (storage == PM_CB_F16 && !(cb_scale > 0.0)) return PM_EINVAL;
  if (sy < nx || sz < sy * ny) return PM_EINVAL;
  const int cx = comp == 0 ? nx : nx - 1;
  const int cy = comp == 1 ? ny : ny - 1;
  const int cz = nz == 1 ? 1 : (comp == 2 ? nz : nz - 1);
  const int64_t n = (int64_t)cx * cy * cz;
  hipLaunchKernelGGL(k_cb, dim3(blocks(n)), dim3(256), 0, (hipStream_t)stream, nx, ny, nz, comp,
                     index, cb_lut, n_lut, storage, cb_scale, sy, sz, out, bad_index);
  return status();
}

}  // extern "C"
