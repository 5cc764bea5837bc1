"""Yee grid description: cell counts, spacings, staggered component extents.

The sizing rules are the reference's (src/prismo/core/grid.py): counts are
``ceil(L * res)`` per axis (grid.py:96-100), the PML layers are added outside the
physical size (grid.py:107-111), a 2-D grid has ``Lz == 0`` with one cell and
``dz == 0`` (grid.py:92-105), and the origin sits ``pml`` cells below zero
(grid.py:113-120).  Component extents follow grid.py:137-176: every component
is one short along each axis on which it sits at a half index.

Device storage is not decided here; the mi355x backend keeps each component in
an x-fastest padded buffer (see ``prismo_amd.backends.mi355x_backend``).
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Tuple, Union

import numpy as np

C0 = 299792458.0

# Half-index offsets of each component on the (dual) Yee cell used by the
# reference: E on cell faces, H on cell edges (grid.py:60-66, solver.py:173-175).
# 1 means "sits at i + 1/2 along that axis" and makes the extent one shorter.
STAGGER = {
    "Ex": (0, 1, 1),
    "Ey": (1, 0, 1),
    "Ez": (1, 1, 0),
    "Hx": (1, 0, 0),
    "Hy": (0, 1, 0),
    "Hz": (0, 0, 1),
}
COMPONENTS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


@dataclass
class GridSpec:
    """Domain size (m), resolution (points per metre) and PML layer count.

    Validation and the scalar-to-triple expansion follow
    src/prismo/core/grid.py:36-50.
    """

    size: Tuple[float, float, float]
    resolution: Union[float, Tuple[float, float, float]]
    boundary_layers: int = 10

    def __post_init__(self):
        if any(s < 0 for s in self.size):
            raise ValueError("Grid size components must be non-negative")
        if isinstance(self.resolution, (int, float)):
            self.resolution = (self.resolution,) * 3
        elif len(self.resolution) != 3:
            raise ValueError("Resolution must be scalar or 3-tuple")
        if any(r <= 0 for r in self.resolution):
            raise ValueError("Resolution must be positive")


class YeeGrid:
    """Grid geometry with the reference's public attributes and helpers."""

    def __init__(self, spec: GridSpec):
        self.spec = spec
        self.Lx, self.Ly, self.Lz = spec.size
        self.res_x, self.res_y, self.res_z = spec.resolution
        self.dx = 1.0 / self.res_x if self.res_x > 0 else 0.0
        self.dy = 1.0 / self.res_y if self.res_y > 0 else 0.0
        self.dz = 1.0 / self.res_z if self.res_z > 0 else 0.0
        self.is_2d = self.Lz == 0.0
        self.is_3d = not self.is_2d

        def count(length, res):
            return int(np.ceil(length * res)) if length > 0 else 1

        self.Nx = count(self.Lx, self.res_x)
        self.Ny = count(self.Ly, self.res_y)
        self.Nz = count(self.Lz, self.res_z) if self.is_3d else 1
        if self.is_2d:
            self.dz = 0.0

        p = spec.boundary_layers
        self.pml_layers = p
        self.Nx_total = self.Nx + 2 * p
        self.Ny_total = self.Ny + 2 * p
        self.Nz_total = self.Nz + (0 if self.is_2d else 2 * p)
        self.origin = np.array(
            [-p * self.dx, -p * self.dy, 0.0 if self.is_2d else -p * self.dz]
        )

    # -- basic properties -------------------------------------------------
    @property
    def dimensions(self) -> Tuple[int, int, int]:
        return (self.Nx_total, self.Ny_total, self.Nz_total)

    @property
    def physical_dimensions(self) -> Tuple[int, int, int]:
        return (self.Nx, self.Ny, self.Nz)

    @property
    def spacing(self) -> Tuple[float, float, float]:
        return (self.dx, self.dy, self.dz)

    def get_field_shape(self, component: str) -> Tuple[int, ...]:
        """Staggered extent of one component (grid.py:137-176)."""
        if component not in STAGGER:
            raise ValueError(f"Unknown field component: {component}")
        sx, sy, sz = STAGGER[component]
        nx, ny, nz = self.dimensions
        if self.is_2d:
            return (nx - sx, ny - sy)
        return (nx - sx, ny - sy, nz - sz)

    def get_coordinates(self, component: str):
        """Physical coordinates of a component's samples (grid.py:178-210)."""
        if component not in STAGGER:
            raise ValueError(f"Unknown field component: {component}")
        sx, sy, sz = STAGGER[component]
        axes = []
        for ax, (n, d, s) in enumerate(
            zip(self.dimensions, self.spacing, (sx, sy, sz))
        ):
            c = self.origin[ax] + np.arange(n) * d
            if s and not (ax == 2 and self.is_2d):
                c = c[:-1] + d / 2
            axes.append(c)
        return tuple(axes[:2]) if self.is_2d else tuple(axes)

    def is_inside_pml(self, i: int, j: int, k: int = 0) -> bool:
        p = self.pml_layers
        if i < p or i >= self.Nx_total - p:
            return True
        if j < p or j >= self.Ny_total - p:
            return True
        return (not self.is_2d) and (k < p or k >= self.Nz_total - p)

    def get_physical_indices(self):
        p = self.pml_layers
        zs = slice(None) if self.is_2d else slice(p, self.Nz_total - p)
        return slice(p, self.Nx_total - p), slice(p, self.Ny_total - p), zs

    # -- time step ----------------------------------------------------------
    def _inv_d2(self) -> float:
        s = (1 / self.dx) ** 2 + (1 / self.dy) ** 2
        if not self.is_2d:
            s += (1 / self.dz) ** 2
        return s

    def get_courant_number(self, dt: float) -> float:
        """S = c dt sqrt(sum 1/d^2) (grid.py:258-284)."""
        return C0 * dt * np.sqrt(self._inv_d2())

    def suggest_time_step(self, safety_factor: float = 0.9) -> float:
        """dt = f / (c sqrt(sum 1/d^2)) (grid.py:304-326)."""
        return safety_factor / (C0 * np.sqrt(self._inv_d2()))

    def get_time_step(self, safety_factor: float = 0.9) -> float:
        return self.suggest_time_step(safety_factor)

    # -- index helpers -------------------------------------------------------
    def point_to_index(self, point) -> Tuple[int, int, int]:
        """Round to the nearest node, then clamp to the physical count minus one
        (grid.py:328-354; note the clamp uses Nx, not Nx_total)."""
        x, y, z = point
        i = int(round((x - self.origin[0]) / self.dx))
        j = int(round((y - self.origin[1]) / self.dy))
        k = int(round((z - self.origin[2]) / self.dz)) if self.is_3d else 0
        i = max(0, min(i, self.Nx - 1))
        j = max(0, min(j, self.Ny - 1))
        k = max(0, min(k, self.Nz - 1))
        return (i, j, k)

    def index_to_coord(self, dim: int, indices):
        if dim not in (0, 1, 2):
            raise ValueError(f"Invalid dimension: {dim}. Must be 0, 1, or 2.")
        return self.origin[dim] + indices * self.spacing[dim]

    def component_box(self, component, x_min, x_max, y_min, y_max, z_min, z_max):
        """Clamped half-open index box of a component region.

        Same clamping as ``get_component_indices`` (grid.py:383-513): the lower
        bounds clamp to [0, N-1] and the upper to [1, N] with the *physical*
        counts, then each staggered axis is cut at the component extent.
        Returns ((i0, i1), (j0, j1), (k0, k1)); k is (0, 1) in 2-D.
        """
        if component not in STAGGER:
            raise ValueError(f"Invalid field component: {component}")
        x_min = max(0, min(x_min, self.Nx - 1))
        x_max = max(1, min(x_max, self.Nx))
        y_min = max(0, min(y_min, self.Ny - 1))
        y_max = max(1, min(y_max, self.Ny))
        if self.is_3d:
            z_min = max(0, min(z_min, self.Nz - 1))
            z_max = max(1, min(z_max, self.Nz))
        else:
            z_min, z_max = 0, 1
        shape = self.get_field_shape(component)
        sx, sy, sz = STAGGER[component]
        i1 = min(x_max, shape[0]) if sx else x_max
        j1 = min(y_max, shape[1]) if sy else y_max
        if self.is_3d:
            k1 = min(z_max, shape[2]) if sz else z_max
        else:
            k1 = z_max
        return (x_min, i1), (y_min, j1), (z_min, k1)

    def get_component_indices(self, component, x_min, x_max, y_min, y_max, z_min, z_max):
        """``np.ix_`` index tuple for a component region (grid.py:383-513)."""
        (i0, i1), (j0, j1), (k0, k1) = self.component_box(
            component, x_min, x_max, y_min, y_max, z_min, z_max
        )
        if self.is_3d:
            return np.ix_(range(i0, i1), range(j0, j1), range(k0, k1))
        return np.ix_(range(i0, i1), range(j0, j1))

    def __repr__(self) -> str:
        kind = "2D" if self.is_2d else "3D"
        return (
            f"YeeGrid({kind}, shape={self.dimensions}, "
            f"spacing=({self.dx:.2e}, {self.dy:.2e}, {self.dz:.2e}), "
            f"PML={self.pml_layers})"
        )


def grid_from_counts(nx: int, ny: int, nz: int, d: float, pml: int = 0) -> YeeGrid:
    """Grid whose *total* node counts are exactly (nx, ny, nz) at spacing d.

    The reference sizes grids by ``ceil(L*res)``, where e.g. 10e-6*50e6 rounds
    up to 501; building from a size a hair under the exact product avoids the
    extra cell.  ``nz == 1`` makes a 2-D grid.
    """
    shrink = 1.0 - 1e-9
    phys = [n - 2 * pml for n in (nx, ny)]
    size_z = 0.0 if nz == 1 else (nz - 2 * pml) * d * shrink
    spec = GridSpec(
        size=(phys[0] * d * shrink, phys[1] * d * shrink, size_z),
        resolution=1.0 / d,
        boundary_layers=pml,
    )
    g = YeeGrid(spec)
    if g.dimensions != (nx, ny, nz):
        raise ValueError(f"could not build grid of {nx}x{ny}x{nz}: got {g.dimensions}")
    return g
