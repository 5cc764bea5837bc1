"""Field monitor (src/prismo/monitors/field.py:18-287).

Time-domain copies of each component region and the running DFT
``F[f] += x * exp(-1j w t) * dt`` (field.py:138-143).  On the device the
regions are sampled into a block of rows and the DFT of a block is one dense
product with the phasor table; results land in the same attributes.
"""

from __future__ import annotations

from typing import Optional, Union

import numpy as np

from .base import Monitor, box_shape, columns_to_region

ALL = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"]


class FieldMonitor(Monitor):
    def __init__(self, center, size, components: Union[list, str] = "all",
                 name: Optional[str] = None, time_domain: bool = True,
                 frequencies: Optional[list] = None):
        super().__init__(center=center, size=size, name=name)
        if components == "all":
            self.components = list(ALL)
        elif components == "E":
            self.components = ["Ex", "Ey", "Ez"]
        elif components == "H":
            self.components = ["Hx", "Hy", "Hz"]
        else:
            for c in components:
                if c not in ALL:
                    raise ValueError(f"Invalid field component: {c}")
            self.components = list(components)
        self.time_domain = time_domain
        self.frequencies = frequencies if frequencies is not None else []
        self._time_data = {c: [] for c in self.components} if time_domain else {}
        self._time_points: list = []
        self._freq_data: dict = {}
        self._freq_stack: dict = {}
        if self.frequencies:
            for c in self.components:
                self._freq_data[c] = {f: None for f in self.frequencies}
        self._component_shapes: dict = {}

    def initialize(self, grid) -> None:
        super().initialize(grid)
        for c in self.components:
            self._component_shapes[c] = tuple(ix.size for ix in self._monitor_region[c])
        self._freq_stack = {}
        if self.frequencies:
            for c in self.components:
                self._set_stack(c, np.zeros((len(self.frequencies),) + self._component_shapes[c],
                                            dtype=np.complex128))

    def _set_stack(self, c, stack) -> None:
        """All frequencies of component c as one (F, ...) array; the
        per-frequency entries are views of it."""
        self._freq_stack[c] = stack
        self._freq_data[c] = {f: stack[fi] for fi, f in enumerate(self.frequencies)}

    def update(self, fields, time: float, dt: float) -> None:
        if self.time_domain:
            self._time_points.append(time)
        for c in self.components:
            data = fields[c][self._monitor_region[c]]
            data = data.detach().cpu().numpy().copy() if hasattr(data, "detach") else data.copy()
            if self.time_domain:
                self._time_data[c].append(data)
            if self.frequencies:
                for f in self.frequencies:
                    omega = 2 * np.pi * f
                    self._freq_data[c][f] += data * np.exp(-1j * omega * time) * dt

    # -- device binding --------------------------------------------------------
    def device_spec(self):
        g = self._grid
        return {
            "boxes": [(c, self._monitor_box[c], box_shape(self._monitor_box[c], g.is_2d))
                      for c in self.components],
            "freqs": np.asarray(self.frequencies, dtype=np.float64) if self.frequencies else None,
            "time_domain": self.time_domain,
        }

    def _ingest_samples(self, times, per_comp) -> None:
        if not self.time_domain:
            return
        self._time_points.extend(float(t) for t in times)
        for c, rows in per_comp.items():
            shape = self._component_shapes[c]
            self._time_data[c].extend(columns_to_region(r, shape) for r in rows)

    def _ingest_spectra(self, per_comp) -> None:
        for c, spec in per_comp.items():
            shape = self._component_shapes[c]
            nf = len(self.frequencies)
            # rows of box-ordered columns (i fastest) -> (F, i, j(, k))
            a = np.asarray(spec, dtype=np.complex128).reshape((nf,) + shape[::-1])
            self._set_stack(c, a.transpose((0,) + tuple(range(len(shape), 0, -1))))

    # -- queries ---------------------------------------------------------------
    def get_time_data(self, component):
        if not self.time_domain:
            raise ValueError("Time domain data not enabled for this monitor")
        if component not in self.components:
            raise ValueError(f"Component {component} not recorded by this monitor")
        return np.array(self._time_points), np.array(self._time_data[component])

    def frequency_stack(self, component) -> np.ndarray:
        """Spectra of ``component`` at every monitor frequency, stacked:
        (F,) + the region shape, in the order of ``frequencies``."""
        if not self.frequencies:
            raise ValueError("Frequency domain data not enabled for this monitor")
        if component not in self.components:
            raise ValueError(f"Component {component} not recorded by this monitor")
        return self._freq_stack[component]

    def get_frequency_data(self, component, frequency):
        if not self.frequencies:
            raise ValueError("Frequency domain data not enabled for this monitor")
        if component not in self.components:
            raise ValueError(f"Component {component} not recorded by this monitor")
        if frequency not in self.frequencies:
            raise ValueError(f"Frequency {frequency} not recorded by this monitor")
        return self._freq_data[component][frequency]

    _PAIRS = [("Ey", "Hz"), ("Ez", "Hy"), ("Ez", "Hx"), ("Ex", "Hz"), ("Ex", "Hy"), ("Ey", "Hx")]
    _SIGNS = [1, -1, 1, -1, 1, -1]

    def get_power_flow(self, frequency=None):
        """Sum of the signed E x H products the reference forms
        (field.py:201-287)."""
        es = [c for c in self.components if c.startswith("E")]
        hs = [c for c in self.components if c.startswith("H")]
        if frequency is None:
            if not self.time_domain:
                raise ValueError("Time domain data not enabled for this monitor")
            if not (es and hs):
                raise ValueError("Need both E and H components to calculate Poynting vector")
            s = 0
            for (e, h), sg in zip(self._PAIRS, self._SIGNS):
                if e in self.components and h in self.components:
                    s = s + sg * self.get_time_data(e)[1] * self.get_time_data(h)[1]
            return np.array(self._time_points), s
        if frequency not in self.frequencies:
            raise ValueError(f"Frequency {frequency} not recorded by this monitor")
        if not (es and hs):
            raise ValueError("Need both E and H components to calculate Poynting vector")
        s = 0
        for (e, h), sg in zip(self._PAIRS, self._SIGNS):
            if e in self.components and h in self.components:
                ed = self.get_frequency_data(e, frequency)
                hd = self.get_frequency_data(h, frequency)
                s = s + 0.5 * sg * np.real(ed * np.conj(hd))
        return s
