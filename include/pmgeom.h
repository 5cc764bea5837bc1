/* prismo_amd geometry -> coefficient build on the device.
 *
 * The reference rasterises each shape on the host (src/prismo/geometry/
 * shapes.py:71-96: np.meshgrid of the coordinate arrays, then the shape's
 * `contains` on every point) and forms the update coefficients from node
 * arrays (src/prismo/core/solver.py:113-133) averaged to each component's
 * points (solver.py:458-533).  These entry points do both on the device:
 * pm_raster writes a material index per grid node (shapes painted in list
 * order, the last one containing a node wins, nodes outside every shape keep
 * index 0), pm_cb_from_index turns the indices into Cb at one E component's
 * sample points through a per-material table of node Cb values.  The
 * containment tests are the reference's expressions in the same FP64
 * operation order, so the indices match its masks bit for bit; the averages
 * keep the reference's summation order.
 *
 * All pointers are device pointers; calls enqueue on `stream` (hipStream_t,
 * NULL = the legacy default stream) and return 0 or a pm_status code.
 */
#ifndef PMGEOM_H
#define PMGEOM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum pm_shape_kind {
  PM_SHAPE_BOX = 1,      /* shapes.py:99-128   |p - c| <= size / 2 on every axis */
  PM_SHAPE_SPHERE = 2,   /* shapes.py:131-157  ||p - c|| <= radius */
  PM_SHAPE_CYLINDER = 3, /* shapes.py:160-212  radial <= radius, |axial| <= height / 2 */
  PM_SHAPE_POLYGON = 4   /* shapes.py:215-282  ray casting in x-y, zmin <= z <= zmax */
};

typedef struct pm_shape {
  int32_t kind;         /* pm_shape_kind */
  int32_t material;     /* index painted where the shape holds a node, 1..255 */
  int32_t axis;         /* cylinder axis, 0..2 */
  int32_t n_vert;       /* polygon vertex count, >= 3 */
  int64_t vert0;        /* polygon: first vertex in `verts` ((x, y) pairs) */
  double center[3];
  double half[3];       /* box: size / 2 per axis (as the reference divides) */
  double radius;        /* sphere, cylinder */
  double half_height;   /* cylinder: height / 2 */
  double zmin, zmax;    /* polygon extrusion */
} pm_shape;

/* out[(k * ny + j) * nx + i] = material index of node (x[i], y[j], z[k])
 * (replaces Shape.rasterize, shapes.py:71-96, for a list of shapes). */
int pm_raster(void* stream, int32_t nx, int32_t ny, int32_t nz, const double* x,
              const double* y, const double* z, const pm_shape* shapes, int32_t n_shapes,
              const double* verts, int64_t n_verts, uint8_t* out);

enum pm_cb_storage { PM_CB_F64 = 0, PM_CB_F32 = 1, PM_CB_F16 = 2 };

/* Cb at the sample points of E component `comp` (0 Ex, 1 Ey, 2 Ez) from the
 * node indices: node value cb_lut[index], averaged as solver.py:458-533
 * (3-D: 0.25 * (((a + b) + c) + d) over the four nodes around each point;
 * 2-D (nz = 1): Ez the same four, Ex / Ey 0.5 * (a + b)), stored into `out`
 * in the padded x-fastest layout (element (i, j, k) at i + j*sy + k*sz) as
 * FP64, FP32, or FP16 holding Cb / cb_scale (rounded to nearest even from
 * FP64, as numpy's float64 -> float16 conversion).  Elements outside the
 * component's index range are not written.  n_lut entries, every index must
 * be below n_lut (checked). */
int pm_cb_from_index(void* stream, int32_t nx, int32_t ny, int32_t nz, const uint8_t* index,
                     const double* cb_lut, int32_t n_lut, int32_t comp, int32_t storage,
                     double cb_scale, int64_t sy, int64_t sz, void* out, int32_t* bad_index);

#ifdef __cplusplus
}
#endif

#endif /* PMGEOM_H */
