"""Monitor base class (src/prismo/monitors/base.py:17-150): the same region
rules as sources, plus the clamped index box of every component."""

from __future__ import annotations

from abc import ABC, abstractmethod
from typing import Optional

import numpy as np


class Monitor(ABC):
    def __init__(self, center, size, name: Optional[str] = None):
        self.center = center
        self.size = size
        self.name = name or f"{self.__class__.__name__}_{id(self)}"
        self._grid = None
        self._monitor_region: dict = {}
        self._monitor_box: dict = {}

    def initialize(self, grid) -> None:
        self._grid = grid
        self._setup_monitor_region()

    def _setup_monitor_region(self) -> None:
        if self._grid is None:
            raise RuntimeError("Monitor must be initialized with a grid first")
        bounds = self._compute_monitor_region()
        for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
            self._monitor_region[comp] = self._grid.get_component_indices(comp, *bounds)
            self._monitor_box[comp] = self._grid.component_box(comp, *bounds)

    def _compute_monitor_region(self):
        if self._grid is None:
            raise RuntimeError("Monitor must be initialized with a grid first")
        hx, hy, hz = self.size[0] / 2, self.size[1] / 2, self.size[2] / 2
        c = self.center
        x0, y0, z0 = self._grid.point_to_index((c[0] - hx, c[1] - hy, c[2] - hz))
        x1, y1, z1 = self._grid.point_to_index((c[0] + hx, c[1] + hy, c[2] + hz))
        if x0 == x1:
            x1 = x0 + 1
        if y0 == y1:
            y1 = y0 + 1
        if z0 == z1 and self._grid.is_3d:
            z1 = z0 + 1
        return x0, x1, y0, y1, z0, z1

    @abstractmethod
    def update(self, fields, time: float, dt: float) -> None: ...

    # -- device binding (overridden by monitors the step library can run) ----
    def device_spec(self):
        """None, or dict(boxes=[(comp, box, shape)], freqs=array|None,
        time_domain=bool) describing what to sample every step."""
        return None

    def _ingest_samples(self, times, per_comp: dict) -> None:  # pragma: no cover
        raise NotImplementedError

    def _ingest_spectra(self, per_comp: dict) -> None:  # pragma: no cover
        raise NotImplementedError


def box_shape(box, is_2d: bool):
    (i0, i1), (j0, j1), (k0, k1) = box
    if is_2d:
        return (i1 - i0, j1 - j0)
    return (i1 - i0, j1 - j0, k1 - k0)


def columns_to_region(v: np.ndarray, shape) -> np.ndarray:
    """Box-ordered (i fastest) values -> region array in [i, j(, k)] layout."""
    return np.asarray(v).reshape(shape, order="F")
