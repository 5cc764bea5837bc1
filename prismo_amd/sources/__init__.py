from .base import Source, SourceTerm
from .gaussian import GaussianBeamSource
from .plane_wave import PlaneWaveSource
from .point import ElectricDipole, MagneticDipole, PointSource
from .waveform import ContinuousWave, CustomWaveform, GaussianPulse, RickerWavelet, Waveform

__all__ = ["Source", "SourceTerm", "Waveform", "ContinuousWave", "GaussianPulse",
           "RickerWavelet", "CustomWaveform", "PointSource", "ElectricDipole",
           "MagneticDipole", "GaussianBeamSource", "PlaneWaveSource"]
