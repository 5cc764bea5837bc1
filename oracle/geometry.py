"""O3 -- CPU restatement of the reference's shape rasterisation.

TEST INFRASTRUCTURE ONLY.  Imported by tests/ to check the device
rasteriser and coefficient build (include/pmgeom.h); never by the product.

Follows src/prismo/geometry/shapes.py: every shape tests points with the
reference's FP64 expressions (Box 122-128, Sphere 153-157, Cylinder 199-212,
Polygon ray casting 243-282) over np.meshgrid(x, y, z, indexing="ij") of the
coordinate arrays (rasterize, 71-96); a list of shapes is painted in order,
the last shape holding a node giving its material index (0 = background).
The node -> component averages of Cb are solver.py:458-533 (restated in
prismo_amd/core/coefficients.py, pinned elsewhere).

Pinned in tests/test_oracle_geometry.py by exact lattice counts that integer
arithmetic fixes: boxes (product of per-axis interval counts), spheres and
cylinders with integer radius on integer lattices (i^2 + j^2 + k^2 <= r^2),
rectangular and L-shaped polygons under the ray-casting rule (half-open
[x0, x1) x [y0, y1) for axis-aligned edges), plus painter's order.
"""

from __future__ import annotations

import numpy as np


def points(x, y, z=None) -> np.ndarray:
    if z is None:
        X, Y = np.meshgrid(x, y, indexing="ij")
        return np.column_stack([X.ravel(), Y.ravel(), np.zeros(X.size)])
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    return np.column_stack([X.ravel(), Y.ravel(), Z.ravel()])


def box_contains(p, center, size):
    half = np.asarray(size, dtype=np.float64) / 2.0
    return np.all(np.abs(p - np.asarray(center, dtype=np.float64)) <= half, axis=1)


def sphere_contains(p, center, radius):
    return np.linalg.norm(p - np.asarray(center, dtype=np.float64), axis=1) <= radius


def cylinder_contains(p, center, radius, height, axis="z"):
    a = {"x": 0, "y": 1, "z": 2}[axis]
    u, v = [i for i in range(3) if i != a]
    c = np.asarray(center, dtype=np.float64)
    axial = np.abs(p[:, a] - c[a])
    radial = np.sqrt((p[:, u] - c[u]) ** 2 + (p[:, v] - c[v]) ** 2)
    return (radial <= radius) & (axial <= height / 2)


def polygon_contains(p, vertices, z_min=-np.inf, z_max=np.inf):
    verts = np.asarray(vertices, dtype=np.float64)
    n = len(verts)
    out = np.zeros(len(p), dtype=bool)
    for idx in range(len(p)):
        xi, yi, zi = p[idx, 0], p[idx, 1], p[idx, 2]
        if not (z_min <= zi <= z_max):
            continue
        count = 0
        for j in range(n):
            v1 = verts[j]
            v2 = verts[(j + 1) % n]
            if ((v1[1] > yi) != (v2[1] > yi)) and (
                    xi < (v2[0] - v1[0]) * (yi - v1[1]) / (v2[1] - v1[1]) + v1[0]):
                count += 1
        out[idx] = count % 2 == 1
    return out


def contains(shape: dict, p) -> np.ndarray:
    """shape: {"kind": "box"|"sphere"|"cylinder"|"polygon", ...} as the
    constructor arguments of the reference's classes."""
    k = shape["kind"]
    if k == "box":
        return box_contains(p, shape["center"], shape["size"])
    if k == "sphere":
        return sphere_contains(p, shape["center"], shape["radius"])
    if k == "cylinder":
        return cylinder_contains(p, shape["center"], shape["radius"], shape["height"],
                                 shape.get("axis", "z"))
    if k == "polygon":
        return polygon_contains(p, shape["vertices"], shape.get("z_min", -np.inf),
                                shape.get("z_max", np.inf))
    raise ValueError(k)


def paint(shapes, x, y, z=None) -> np.ndarray:
    """Material index per node, shape (nx, ny, nz) (nz = 1 in 2-D): each
    shape's ``material`` where it holds the node, later shapes on top."""
    p = points(x, y, z)
    out = np.zeros(len(p), dtype=np.uint8)
    for s in shapes:
        out[contains(s, p)] = s["material"]
    nz = 1 if z is None else len(z)
    return out.reshape(len(x), len(y), nz)
