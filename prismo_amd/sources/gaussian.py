"""Gaussian-beam source (src/prismo/sources/gaussian.py:18-298).

The transverse profile exp(-r^2 / w0^2) is evaluated on the unclamped source
region at the source plane (gaussian.py:188-256); the per-step value is
amplitude * profile for E and that divided by 377 for H, added with NumPy's
broadcasting of ``values.flat`` against the component region
(gaussian.py:291-298).  The device term stores the broadcast profile once and
the amplitude per step, computing the same products.
"""

from __future__ import annotations

from typing import Optional

import numpy as np

from .base import Source, SourceTerm
from .plane_wave import COMPONENT_MAP
from .point import _make_waveform


class GaussianBeamSource(Source):
    def __init__(self, center, size, direction: str, polarization: str, frequency: float,
                 beam_waist: float, pulse: bool = True, pulse_width: Optional[float] = None,
                 amplitude: float = 1.0, phase: float = 0.0, name: Optional[str] = None,
                 enabled: bool = True):
        super().__init__(center=center, size=size, name=name, enabled=enabled)
        if direction.lower() == polarization.lower():
            raise ValueError(
                f"Polarization ({polarization}) must be perpendicular to "
                f"propagation direction ({direction})"
            )
        self.direction = direction
        self.polarization = polarization
        self.frequency = frequency
        self.beam_waist = beam_waist
        self.wavelength = 299792458.0 / frequency
        self.k = 2 * np.pi / self.wavelength
        self.rayleigh_range = np.pi * beam_waist**2 / self.wavelength
        self.waveform = _make_waveform(frequency, pulse, pulse_width, amplitude, phase)
        self._e_components: list = []
        self._h_components: list = []

    def initialize(self, grid) -> None:
        super().initialize(grid)
        key = (self.direction.lower(), self.polarization.lower())
        if key not in COMPONENT_MAP:
            raise ValueError(
                f"Invalid direction ({self.direction}) and polarization "
                f"({self.polarization}) combination"
            )
        self._e_components = COMPONENT_MAP[key]["E"]
        self._h_components = COMPONENT_MAP[key]["H"]

    def _profile(self) -> np.ndarray:
        """exp(-r^2 / w0^2) on the source plane, shape as the reference builds it."""
        g = self._grid
        x0, x1, y0, y1, z0, z1 = self._compute_source_region()
        if not g.is_3d and z0 == z1:
            z1 = z0 + 1
        d = self.direction.lower()
        if d == "x":
            ys, zs = np.meshgrid(np.arange(y0, y1), np.arange(z0, z1), indexing="ij")
            yp = g.index_to_coord(1, ys)
            if g.is_3d:
                r2 = (yp - self.center[1]) ** 2 + (g.index_to_coord(2, zs) - self.center[2]) ** 2
            else:
                r2 = (yp - self.center[1]) ** 2
        elif d == "y":
            xs, zs = np.meshgrid(np.arange(x0, x1), np.arange(z0, z1), indexing="ij")
            xp = g.index_to_coord(0, xs)
            if g.is_3d:
                r2 = (xp - self.center[0]) ** 2 + (g.index_to_coord(2, zs) - self.center[2]) ** 2
            else:
                r2 = (xp - self.center[0]) ** 2
        else:
            xs, ys = np.meshgrid(np.arange(x0, x1), np.arange(y0, y1), indexing="ij")
            r2 = (g.index_to_coord(0, xs) - self.center[0]) ** 2 + (
                g.index_to_coord(1, ys) - self.center[1]) ** 2
        return np.exp(-r2 / (self.beam_waist**2))

    def _calculate_beam_profile(self, time: float) -> dict:
        prof = self.waveform(time) * self._profile()
        out = {c: prof for c in self._e_components}
        out.update({c: prof / 377.0 for c in self._h_components})
        return out

    def update_fields(self, fields, time: float, dt: float) -> None:
        if not self.enabled or self._grid is None:
            return
        import torch

        for comp, values in self._calculate_beam_profile(time).items():
            idx = self._source_region[comp]
            target = fields[comp]
            vals = np.asarray(values.flat)
            if isinstance(target, torch.Tensor):
                cur = target[idx]
                target[idx] = cur + torch.as_tensor(vals, dtype=cur.dtype, device=cur.device)
            else:
                target[idx] += vals

    def device_terms(self, times):
        if not self.enabled or self._grid is None:
            return []
        flat = np.asarray(self._profile().flat)
        amp = self.waveform.table(times)
        terms = []
        for comp in self._e_components + self._h_components:
            box = self._source_box[comp]
            (i0, i1), (j0, j1), (k0, k1) = box
            shape = (i1 - i0, j1 - j0) if not self._grid.is_3d else (i1 - i0, j1 - j0, k1 - k0)
            # NumPy broadcasting of the flat profile against the region (raises
            # exactly where the reference's in-place add would)
            prof = np.broadcast_to(flat, shape)
            prof_box = np.asfortranarray(prof).ravel(order="F")
            terms.append(SourceTerm(comp, box, amp, prof_box.astype(np.float64),
                                    377.0 if comp in self._h_components else 0.0))
        return terms
