"""Plane-wave (line / sheet) source (src/prismo/sources/plane_wave.py:16-221):
E += a * sign and H += (a / 377) * sign on the direction/polarisation pair of
components."""

from __future__ import annotations

from typing import Optional

import numpy as np

from .base import Source, SourceTerm
from .point import _make_waveform

COMPONENT_MAP = {
    ("x", "y"): {"E": ["Ey"], "H": ["Hz"]},
    ("x", "z"): {"E": ["Ez"], "H": ["Hy"]},
    ("y", "x"): {"E": ["Ex"], "H": ["Hz"]},
    ("y", "z"): {"E": ["Ez"], "H": ["Hx"]},
    ("z", "x"): {"E": ["Ex"], "H": ["Hy"]},
    ("z", "y"): {"E": ["Ey"], "H": ["Hx"]},
}


class PlaneWaveSource(Source):
    def __init__(self, center, size, direction: str, polarization: str, frequency: float,
                 pulse: bool = True, pulse_width: Optional[float] = None,
                 amplitude: float = 1.0, phase: float = 0.0, name: Optional[str] = None,
                 enabled: bool = True):
        super().__init__(center=center, size=size, name=name, enabled=enabled)
        if direction.startswith("+"):
            self.direction, self.direction_sign = direction[1:], 1
        elif direction.startswith("-"):
            self.direction, self.direction_sign = direction[1:], -1
        else:
            self.direction, self.direction_sign = direction, 1
        if self.direction.lower() not in ("x", "y", "z"):
            raise ValueError(
                f"Invalid direction: {direction}. Must be one of: x, y, z, +x, -x, +y, -y, +z, -z"
            )
        if self.direction.lower() == polarization.lower():
            raise ValueError(
                f"Polarization ({polarization}) must be perpendicular to "
                f"propagation direction ({self.direction})"
            )
        self.polarization = polarization
        self.frequency = frequency
        self.wavelength = 299792458.0 / frequency
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

    def update_fields(self, fields, time: float, dt: float) -> None:
        if not self.enabled or self._grid is None:
            return
        a = self.waveform(time)
        for c in self._e_components:
            fields[c][self._source_region[c]] += float(a * self.direction_sign)
        for c in self._h_components:
            fields[c][self._source_region[c]] += float((a / 377.0) * self.direction_sign)

    def device_terms(self, times):
        if not self.enabled or self._grid is None:
            return []
        amp = self.waveform.table(times)
        s = float(self.direction_sign)
        # elementwise IEEE operations: the same bits as the per-step scalars
        terms = [SourceTerm(c, self._source_box[c], amp * s) for c in self._e_components]
        terms += [SourceTerm(c, self._source_box[c], (amp / 377.0) * s)
                  for c in self._h_components]
        return terms
