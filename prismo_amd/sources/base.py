"""Source base class (src/prismo/sources/base.py:17-165).

Region rules are the reference's: the centre +- size/2 corners go through
``point_to_index`` (round, then clamp to the physical count minus one), an
empty extent is widened to one cell, and each component keeps its own
clamped ``np.ix_`` region.  Sources that can run on the device also describe
themselves as ``SourceTerm``s: a component box plus one value per step
(optionally times a per-cell profile), which the step library adds after the
E update.  A subclass without ``device_terms`` still works: the simulation
calls its ``update_fields`` on the device field views every step.
"""

from __future__ import annotations

from abc import ABC, abstractmethod
from dataclasses import dataclass, field
from typing import Optional

import numpy as np


@dataclass
class SourceTerm:
    comp: str
    box: tuple                     # ((i0, i1), (j0, j1), (k0, k1))
    table: np.ndarray              # FP64 value per step
    profile: Optional[np.ndarray] = None  # per cell, box order i fastest
    divisor: float = 0.0


class Source(ABC):
    def __init__(self, center, size, name: Optional[str] = None, enabled: bool = True):
        self.center = center
        self.size = size
        self.name = name or f"{self.__class__.__name__}_{id(self)}"
        self.enabled = enabled
        self._grid = None
        self._source_region: dict = {}
        self._source_box: dict = {}

    def initialize(self, grid) -> None:
        self._grid = grid
        self._setup_source_region()

    def _setup_source_region(self) -> None:
        if self._grid is None:
            raise RuntimeError("Source must be initialized with a grid first")
        bounds = self._compute_source_region()
        self._source_region = {}
        self._source_box = {}
        for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
            self._source_region[comp] = self._grid.get_component_indices(comp, *bounds)
            self._source_box[comp] = self._grid.component_box(comp, *bounds)

    def _compute_source_region(self):
        if self._grid is None:
            raise RuntimeError("Source must be initialized with a grid first")
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
    def update_fields(self, fields, time: float, dt: float) -> None: ...

    def device_terms(self, times) -> Optional[list]:
        """Per-step additive terms for ``times`` (None: not expressible)."""
        return None

    def disable(self) -> None:
        self.enabled = False

    def enable(self) -> None:
        self.enabled = True


def box_cells(box) -> int:
    (i0, i1), (j0, j1), (k0, k1) = box
    return max(0, i1 - i0) * max(0, j1 - j0) * max(0, k1 - k0)
