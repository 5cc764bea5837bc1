"""Geometric shapes (src/prismo/geometry/shapes.py:15-376).

Same classes, constructor arguments and host ``contains`` / ``rasterize``
semantics as the reference (FP64, closed bounds, ray casting for polygons).
Each analytic shape also describes itself as a ``pm_shape`` record
(include/pmgeom.h) so a list of shapes can be painted onto the grid on the
device (``prismo_amd.geometry.raster``) instead of through host meshgrids.
"""

from __future__ import annotations

from abc import ABC, abstractmethod
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np


@dataclass
class Material:
    """Material of a shape (shapes.py:15-33)."""

    name: str
    epsilon_r: float = 1.0
    mu_r: float = 1.0


class Shape(ABC):
    def __init__(self, material: Material, center):
        self.material = material
        self.center = np.array(center)

    @abstractmethod
    def contains(self, points: np.ndarray) -> np.ndarray: ...

    def rasterize(self, x, y, z=None) -> np.ndarray:
        """Boolean mask over np.meshgrid(x, y[, z], indexing="ij")
        (shapes.py:71-96; 2-D points carry z = 0)."""
        if z is None:
            X, Y = np.meshgrid(x, y, indexing="ij")
            pts = np.column_stack([X.ravel(), Y.ravel(), np.zeros(X.size)])
            return self.contains(pts).reshape(X.shape)
        X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
        pts = np.column_stack([X.ravel(), Y.ravel(), Z.ravel()])
        return self.contains(pts).reshape(X.shape)

    def device_record(self) -> Optional[dict]:
        """Fields of the pm_shape record (None: no device form)."""
        return None


class Box(Shape):
    def __init__(self, material: Material, center, size):
        super().__init__(material, center)
        self.size = np.array(size)
        self.half_size = self.size / 2.0

    def contains(self, points):
        return np.all(np.abs(points - self.center) <= self.half_size, axis=1)

    def device_record(self):
        return {"kind": 1, "center": self.center, "half": self.half_size}


class Sphere(Shape):
    def __init__(self, material: Material, center, radius: float):
        super().__init__(material, center)
        self.radius = radius

    def contains(self, points):
        return np.linalg.norm(points - self.center, axis=1) <= self.radius

    def device_record(self):
        return {"kind": 2, "center": self.center, "radius": float(self.radius)}


class Cylinder(Shape):
    def __init__(self, material: Material, center, radius: float, height: float,
                 axis: str = "z"):
        super().__init__(material, center)
        self.radius = radius
        self.height = height
        self.axis = axis.lower()
        if self.axis not in ["x", "y", "z"]:
            raise ValueError("axis must be 'x', 'y', or 'z'")

    def contains(self, points):
        a = {"x": 0, "y": 1, "z": 2}[self.axis]
        u, v = [i for i in range(3) if i != a]
        axial = np.abs(points[:, a] - self.center[a])
        radial = np.sqrt((points[:, u] - self.center[u]) ** 2
                         + (points[:, v] - self.center[v]) ** 2)
        return (radial <= self.radius) & (axial <= self.height / 2)

    def device_record(self):
        return {"kind": 3, "center": self.center, "radius": float(self.radius),
                "half_height": self.height / 2, "axis": {"x": 0, "y": 1, "z": 2}[self.axis]}


class Polygon(Shape):
    """Polygon in x-y, extruded over [z_min, z_max] (ray casting)."""

    def __init__(self, material: Material, vertices, z_min: float = -np.inf,
                 z_max: float = np.inf):
        vertices = np.asarray(vertices)
        c = np.mean(vertices, axis=0)
        super().__init__(material, np.array([c[0], c[1], (z_min + z_max) / 2]))
        self.vertices = vertices
        self.z_min = z_min
        self.z_max = z_max

    def contains(self, points):
        x, y = points[:, 0], points[:, 1]
        z = points[:, 2] if points.shape[1] > 2 else np.zeros_like(x)
        zin = (z >= self.z_min) & (z <= self.z_max)
        v = np.asarray(self.vertices, dtype=np.float64)
        w = np.roll(v, -1, axis=0)
        count = np.zeros(len(points), dtype=np.int64)
        for (x1, y1), (x2, y2) in zip(v, w):
            cross = (y1 > y) != (y2 > y)
            with np.errstate(divide="ignore", invalid="ignore"):
                xc = (x2 - x1) * (y - y1) / (y2 - y1) + x1
            count += cross & (x < xc)
        return (count % 2 == 1) & zin

    def device_record(self):
        if len(self.vertices) < 3:
            return None
        return {"kind": 4, "vertices": np.asarray(self.vertices, dtype=np.float64),
                "zmin": float(self.z_min), "zmax": float(self.z_max)}


class CustomShape(Shape):
    """Shape given by a host function of the (N, 3) points (no device form)."""

    def __init__(self, material: Material, center, function: Callable[[np.ndarray], np.ndarray]):
        super().__init__(material, center)
        self.function = function

    def contains(self, points):
        return self.function(points)


class GeometryGroup:
    """Boolean combination of shapes (shapes.py:316-376)."""

    def __init__(self, shapes: list, operation: str = "union"):
        self.shapes = shapes
        self.operation = operation.lower()
        if self.operation not in ["union", "intersection", "difference"]:
            raise ValueError("operation must be 'union', 'intersection', or 'difference'")

    def rasterize(self, x, y, z=None):
        if len(self.shapes) == 0:
            shape = (len(x), len(y), len(z) if z is not None else 1)
            return np.zeros(shape, dtype=bool), []
        masks = [s.rasterize(x, y, z) for s in self.shapes]
        combined = masks[0]
        for m in masks[1:]:
            if self.operation == "union":
                combined = combined | m
            elif self.operation == "intersection":
                combined = combined & m
            else:
                combined = combined & ~m
        return combined, masks
