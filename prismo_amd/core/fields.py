"""The six field components (src/prismo/core/fields.py:20-415).

Same interface: dict-like access, in-place ``__setitem__`` with shape check,
energy, magnitude, PEC helper, memory report.  On the mi355x backend each
component is a reference-shaped view of a padded x-fastest device buffer
(``MI355XBackend.field_zeros``), which is what the step library updates.
"""

from __future__ import annotations

from typing import Optional

import numpy as np

from ..backends.backend_manager import resolve_backend
from .grid import COMPONENTS

EPS0 = 8.854187817e-12
MU0 = 4 * np.pi * 1e-7


def _is_torch(a) -> bool:
    return hasattr(a, "detach") and hasattr(a, "data_ptr")


class ElectromagneticFields:
    def __init__(self, grid, dtype=None, backend=None):
        self.grid = grid
        self.backend = resolve_backend(backend)
        self.dtype = self.backend.float64 if dtype is None else dtype
        self._fields: dict = {}
        self._initialize_fields()

    def _initialize_fields(self) -> None:
        b = self.backend
        if hasattr(b, "field_set"):
            self._fields.update(b.field_set(self.grid, self.dtype))
            return
        for c in COMPONENTS:
            if hasattr(b, "field_zeros"):
                self._fields[c] = b.field_zeros(self.grid, c, self.dtype)
            else:
                self._fields[c] = b.zeros(self.grid.get_field_shape(c), dtype=self.dtype)

    def __getitem__(self, component):
        if component not in self._fields:
            raise KeyError(f"Unknown field component: {component}")
        return self._fields[component]

    def __setitem__(self, component, value) -> None:
        if component not in self._fields:
            raise KeyError(f"Unknown field component: {component}")
        tgt = self._fields[component]
        if np.isscalar(value):
            tgt.fill_(value) if _is_torch(tgt) else tgt.fill(value)
            return
        arr = self.backend.asarray(value)
        if tuple(arr.shape) != tuple(tgt.shape):
            raise ValueError(
                f"Shape mismatch for {component}: expected {tuple(tgt.shape)}, "
                f"got {tuple(arr.shape)}"
            )
        if _is_torch(tgt):
            tgt.copy_(arr)
        else:
            tgt[...] = arr

    Ex = property(lambda self: self._fields["Ex"])
    Ey = property(lambda self: self._fields["Ey"])
    Ez = property(lambda self: self._fields["Ez"])
    Hx = property(lambda self: self._fields["Hx"])
    Hy = property(lambda self: self._fields["Hy"])
    Hz = property(lambda self: self._fields["Hz"])

    def get_electric_field_components(self):
        return self.Ex, self.Ey, self.Ez

    def get_magnetic_field_components(self):
        return self.Hx, self.Hy, self.Hz

    def get_field_type_components(self, field_type):
        if field_type == "E":
            return self.get_electric_field_components()
        if field_type == "H":
            return self.get_magnetic_field_components()
        raise ValueError(f"Unknown field type: {field_type}")

    def zero_fields(self, field_type: Optional[str] = None) -> None:
        if field_type is None:
            comps = COMPONENTS
        elif field_type == "E":
            comps = ("Ex", "Ey", "Ez")
        elif field_type == "H":
            comps = ("Hx", "Hy", "Hz")
        else:
            raise ValueError(f"Unknown field type: {field_type}")
        for c in comps:
            f = self._fields[c]
            f.zero_() if _is_torch(f) else f.fill(0.0)

    def copy_fields_from(self, other: "ElectromagneticFields") -> None:
        if not isinstance(other, ElectromagneticFields):
            raise TypeError("Can only copy from another ElectromagneticFields object")
        if self.grid.dimensions != other.grid.dimensions:
            raise ValueError("Grid dimensions must match for field copying")
        for c in self._fields:
            if tuple(self._fields[c].shape) != tuple(other._fields[c].shape):
                raise ValueError(f"Shape mismatch for component {c}")
            self[c] = other._fields[c]

    def to_numpy(self, component: str) -> np.ndarray:
        return self.backend.to_numpy(self._fields[component])

    def _sum_sq(self, a) -> float:
        if _is_torch(a):
            return float((a.double() ** 2).sum().item())
        return float(np.sum(a**2))

    def get_field_energy(self, field_type: Optional[str] = None, region=None) -> float:
        """0.5 (eps0 |E|^2 + mu0 |H|^2) dV summed (fields.py:237-285)."""
        dx, dy, dz = self.grid.spacing
        dV = dx * dy if self.grid.is_2d else dx * dy * dz
        total = 0.0
        if field_type in (None, "E"):
            for c in ("Ex", "Ey", "Ez"):
                f = self._fields[c] if region is None else self._fields[c][region]
                total += 0.5 * EPS0 * self._sum_sq(f) * dV
        if field_type in (None, "H"):
            for c in ("Hx", "Hy", "Hz"):
                f = self._fields[c] if region is None else self._fields[c][region]
                total += 0.5 * MU0 * self._sum_sq(f) * dV
        return total

    def get_field_magnitude(self, field_type, region=None):
        fx, fy, fz = (self.backend.to_numpy(f) for f in self.get_field_type_components(field_type))
        if region is not None:
            fx, fy, fz = fx[region], fy[region], fz[region]
        m = tuple(min(s) for s in zip(fx.shape, fy.shape, fz.shape))
        cut = tuple(slice(0, n) for n in m)
        return np.sqrt(fx[cut] ** 2 + fy[cut] ** 2 + fz[cut] ** 2)

    def apply_boundary_conditions(self, boundary_type: str = "pec") -> None:
        if boundary_type.lower() != "pec":
            raise NotImplementedError(f"Boundary condition '{boundary_type}' not implemented")
        self.Ey[0, :] = 0
        self.Ey[-1, :] = 0
        self.Ez[0, :] = 0
        self.Ez[-1, :] = 0
        self.Ex[:, 0] = 0
        self.Ex[:, -1] = 0
        self.Ez[:, 0] = 0
        self.Ez[:, -1] = 0
        if not self.grid.is_2d:
            self.Ex[:, :, 0] = 0
            self.Ex[:, :, -1] = 0
            self.Ey[:, :, 0] = 0
            self.Ey[:, :, -1] = 0

    def get_memory_usage(self) -> dict:
        total = 0
        info = {}
        for c, f in self._fields.items():
            nbytes = int(np.prod(f.shape)) * (f.element_size() if _is_torch(f) else f.itemsize)
            total += nbytes
            info[c] = {"shape": tuple(f.shape), "dtype": str(f.dtype), "bytes": nbytes,
                       "megabytes": nbytes / 2**20}
        return {"total_bytes": total, "total_megabytes": total / 2**20,
                "total_gigabytes": total / 2**30, "components": info}

    def __repr__(self) -> str:
        mi = self.get_memory_usage()
        return (f"ElectromagneticFields(grid={self.grid.dimensions}, "
                f"dtype={self.dtype}, memory={mi['total_megabytes']:.1f}MB)")

    __str__ = __repr__
