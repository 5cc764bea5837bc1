"""On-the-fly DFT monitor (API of src/prismo/monitors/dft.py:18-273).

The reference accumulates ``field[:10, :10]`` whatever the region
(dft.py:92-106, 139-160).  Here the accumulators cover the monitor region of
each component (the FieldMonitor semantics of field.py:138-143);
``compat_placeholder=True`` restores the reference's fixed [:10, :10] corner
(2-D) so its numbers can be reproduced.
"""

from __future__ import annotations

from typing import Optional

import numpy as np

from ..backends.backend_manager import resolve_backend
from .base import Monitor, box_shape, columns_to_region


class DFTMonitor(Monitor):
    def __init__(self, center, size, frequencies, components: Optional[list] = None,
                 name: Optional[str] = None, backend=None, compat_placeholder: bool = False):
        super().__init__(center, size, name)
        self.frequencies = np.array(frequencies)
        self.omega = 2 * np.pi * self.frequencies
        self.components = ["Ex", "Ey", "Ez"] if components is None else list(components)
        self.backend = resolve_backend(backend)
        self.compat_placeholder = compat_placeholder
        self._dft_data: dict = {}
        self._time_steps = 0
        self._dt = None
        self._boxes: dict = {}

    def initialize(self, grid) -> None:
        super().initialize(grid)
        for c in self.components:
            self._boxes[c] = self._component_box(c)
            shape = box_shape(self._boxes[c], grid.is_2d)
            self._dft_data[c] = np.zeros((len(self.frequencies),) + shape, dtype=np.complex128)

    def _component_box(self, c):
        g = self._grid
        if not self.compat_placeholder:
            return self._monitor_box[c]
        if g.is_3d:
            raise ValueError("the [:10, :10] placeholder region does not broadcast in 3-D")
        ext = g.get_field_shape(c)
        return ((0, min(10, ext[0])), (0, min(10, ext[1])), (0, 1))

    def _get_component_shape(self, component):
        if self._grid is None:
            raise RuntimeError("Monitor must be initialized first")
        return box_shape(self._boxes[component], self._grid.is_2d)

    def _extract_field_data(self, fields, component):
        (i0, i1), (j0, j1), (k0, k1) = self._boxes[component]
        f = fields[component]
        sub = f[i0:i1, j0:j1] if self._grid.is_2d else f[i0:i1, j0:j1, k0:k1]
        return sub.detach().cpu().numpy() if hasattr(sub, "detach") else np.asarray(sub)

    def update(self, fields, time: float, dt: float) -> None:
        if self._dt is None:
            self._dt = dt
        for c in self.components:
            data = self._extract_field_data(fields, c)
            for fi, omega in enumerate(self.omega):
                self._dft_data[c][fi] += data * np.exp(-1j * omega * time) * dt
        self._time_steps += 1

    # -- device binding --------------------------------------------------------
    def device_spec(self):
        g = self._grid
        return {
            "boxes": [(c, self._boxes[c], box_shape(self._boxes[c], g.is_2d))
                      for c in self.components],
            "freqs": np.asarray(self.frequencies, dtype=np.float64),
            "time_domain": False,
        }

    def _ingest_samples(self, times, per_comp) -> None:
        self._time_steps += len(times)

    def _note_dt(self, dt):
        if self._dt is None:
            self._dt = dt

    def _ingest_spectra(self, per_comp) -> None:
        for c, spec in per_comp.items():
            shape = self._get_component_shape(c)
            self._dft_data[c] = np.stack(
                [columns_to_region(spec[fi], shape) for fi in range(len(self.frequencies))]
            ).astype(np.complex128)

    # -- queries ---------------------------------------------------------------
    def get_frequency_data(self, component, frequency_index=None):
        if component not in self._dft_data:
            raise ValueError(f"Component {component} not monitored")
        if frequency_index is None:
            return {f: self._dft_data[component][i] for i, f in enumerate(self.frequencies)}
        return self._dft_data[component][frequency_index]

    def get_intensity(self, component, frequency_index):
        return np.abs(self.get_frequency_data(component, frequency_index)) ** 2

    def get_power_spectrum(self, component, normalize: bool = True):
        s = np.array([np.sum(self.get_intensity(component, i))
                      for i in range(len(self.frequencies))])
        if normalize and np.max(s) > 0:
            s = s / np.max(s)
        return s

    def get_transmission_spectrum(self, reference_power=None):
        p = np.zeros(len(self.frequencies))
        for i in range(len(self.frequencies)):
            for c in ("Ex", "Ey", "Ez"):
                if c in self._dft_data:
                    p[i] += np.sum(self.get_intensity(c, i))
        if reference_power is not None:
            return p / reference_power
        if np.max(p) > 0:
            return p / np.max(p)
        return p
