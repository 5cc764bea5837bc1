"""Geometry -> coefficients on the device (include/pmgeom.h).

``paint`` turns a list of shapes into a material index per grid node with one
kernel (painter's order: later shapes cover earlier ones, 0 = background);
``cb_from_index`` forms Cb at one E component's sample points from those
indices and a per-material table of node Cb values, straight into the step
library's padded coefficient arrays.  Together they replace the host path of
node arrays -> cell_coefficients -> component_average -> upload
(src/prismo/core/solver.py:64-133, 458-533) for lossless, non-magnetic
structures, with the same bits.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _native as N


def _torch():
    import torch

    return torch


def material_table(shapes, background=None) -> list:
    """Distinct materials in order of first use; index 0 is ``background``."""
    mats = [background]
    for s in shapes:
        if not any(s.material is m for m in mats):
            mats.append(s.material)
    if len(mats) > 256:
        raise ValueError("at most 255 distinct shape materials")
    return mats


def shape_records(shapes, mats):
    """(pm_shape array, packed polygon vertices) for ``shapes``."""
    recs = (N.pm_shape * max(1, len(shapes)))()
    verts = []
    nv = 0
    for t, s in enumerate(shapes):
        d = s.device_record()
        if d is None:
            raise NotImplementedError(
                f"{type(s).__name__} has no device form; give the structure as "
                "material_arrays instead")
        r = recs[t]
        r.kind = d["kind"]
        r.material = next(i for i, m in enumerate(mats) if m is s.material)
        c = np.asarray(d.get("center", (0.0, 0.0, 0.0)), dtype=np.float64)
        for a in range(3):
            r.center[a] = float(c[a]) if a < len(c) else 0.0
        h = np.asarray(d.get("half", (0.0, 0.0, 0.0)), dtype=np.float64)
        for a in range(3):
            r.half[a] = float(h[a]) if a < len(h) else 0.0
        r.radius = d.get("radius", 0.0)
        r.half_height = d.get("half_height", 0.0)
        r.axis = d.get("axis", 2)
        r.zmin = d.get("zmin", -np.inf)
        r.zmax = d.get("zmax", np.inf)
        if d["kind"] == 4:
            v = d["vertices"]
            r.vert0 = nv
            r.n_vert = len(v)
            verts.append(v[:, :2].reshape(-1))
            nv += len(v)
    packed = np.concatenate(verts) if verts else np.zeros(2, dtype=np.float64)
    return recs, packed, nv


def paint(shapes, x, y, z, device, mats=None):
    """Material index per node as a (nz, ny, nx) uint8 device tensor
    (2-D: z = [0.0]) and the material table."""
    torch = _torch()
    lib = N.load()
    mats = mats if mats is not None else material_table(shapes)
    recs, verts, nv = shape_records(shapes, mats)
    nx, ny, nz = len(x), len(y), len(z)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(device)
    tx, ty, tz, tv = dev(x), dev(y), dev(z), dev(verts)
    traw = torch.frombuffer(bytearray(bytes(recs)), dtype=torch.uint8).to(device)
    out = torch.empty((nz, ny, nx), dtype=torch.uint8, device=device)
    stream = torch.cuda.current_stream(device).cuda_stream
    rc = lib.pm_raster(stream, nx, ny, nz, tx.data_ptr(), ty.data_ptr(), tz.data_ptr(),
                       traw.data_ptr(), len(shapes), tv.data_ptr(), nv, out.data_ptr())
    if rc != 0:
        raise N.NativeError(f"pm_raster failed ({rc})")
    # keep the inputs alive until the kernel has read them
    torch.cuda.current_stream(device).synchronize()
    return out, mats


def cb_from_index(idx, cb_lut, comp: int, storage: int, cb_scale: float, out, sy: int, sz: int):
    """Cb of E component ``comp`` into ``out`` (padded layout) from the node
    indices ``idx`` ((nz, ny, nx) uint8) and the per-material node Cb."""
    torch = _torch()
    lib = N.load()
    nz, ny, nx = idx.shape
    lut = torch.from_numpy(np.ascontiguousarray(cb_lut, dtype=np.float64)).to(idx.device)
    bad = torch.zeros(1, dtype=torch.int32, device=idx.device)
    stream = torch.cuda.current_stream(idx.device).cuda_stream
    rc = lib.pm_cb_from_index(stream, nx, ny, nz, idx.data_ptr(), lut.data_ptr(), len(cb_lut),
                              comp, storage, float(cb_scale), sy, sz, out.data_ptr(),
                              bad.data_ptr())
    if rc != 0:
        raise N.NativeError(f"pm_cb_from_index failed ({rc})")
    if int(bad.item()) != 0:
        raise ValueError("material index beyond the table")
