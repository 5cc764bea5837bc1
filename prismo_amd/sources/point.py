"""Point dipole sources (src/prismo/sources/point.py:15-205):
``field[region] += waveform(t)`` on one component."""

from __future__ import annotations

from typing import Optional

from .base import Source, SourceTerm
from .waveform import ContinuousWave, GaussianPulse, Waveform


class PointSource(Source):
    def __init__(self, position, component: str, waveform: Waveform,
                 name: Optional[str] = None, enabled: bool = True):
        super().__init__(center=position, size=(0, 0, 0), name=name, enabled=enabled)
        self.component = component
        self.waveform = waveform

    def update_fields(self, fields, time: float, dt: float) -> None:
        if not self.enabled:
            return
        amplitude = self.waveform(time)
        fields[self.component][self._source_region[self.component]] += float(amplitude)

    def device_terms(self, times):
        if not self.enabled:
            return []
        return [SourceTerm(self.component, self._source_box[self.component],
                           self.waveform.table(times))]


def _make_waveform(frequency, pulse, pulse_width, amplitude, phase):
    if pulse:
        if pulse_width is None:
            raise ValueError("pulse_width must be provided for pulsed sources")
        return GaussianPulse(frequency=frequency, pulse_width=pulse_width,
                             amplitude=amplitude, phase=phase)
    return ContinuousWave(frequency=frequency, amplitude=amplitude, phase=phase)


class ElectricDipole(PointSource):
    def __init__(self, position, polarization: str, frequency: float, pulse: bool = True,
                 pulse_width: Optional[float] = None, amplitude: float = 1.0,
                 phase: float = 0.0, name: Optional[str] = None, enabled: bool = True):
        comp = {"x": "Ex", "y": "Ey", "z": "Ez"}[polarization.lower()]
        super().__init__(position, comp,
                         _make_waveform(frequency, pulse, pulse_width, amplitude, phase),
                         name=name, enabled=enabled)


class MagneticDipole(PointSource):
    def __init__(self, position, polarization: str, frequency: float, pulse: bool = True,
                 pulse_width: Optional[float] = None, amplitude: float = 1.0,
                 phase: float = 0.0, name: Optional[str] = None, enabled: bool = True):
        comp = {"x": "Hx", "y": "Hy", "z": "Hz"}[polarization.lower()]
        super().__init__(position, comp,
                         _make_waveform(frequency, pulse, pulse_width, amplitude, phase),
                         name=name, enabled=enabled)
