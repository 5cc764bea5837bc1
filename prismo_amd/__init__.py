"""prismo_amd: an MI355X (gfx950) FDTD time-step core behind prismo's
backend interface.

``prismo_amd.set_backend('mi355x')`` selects torch-ROCm device arrays; the
time step (Yee E/H updates, CPML, ADE, sources, monitor DFT) runs in the
hand-written HIP library ``libpmfdtd.so`` through a C ABI.  Names mirror
the reference package (src/prismo/__init__.py) for the hot path.
"""

from . import backends, boundaries, core, geometry, materials, modes, monitors, sources
from .backends import Backend, get_backend, list_available_backends, set_backend
from .boundaries.pml import CPML, PMLParams
from .core.fields import ElectromagneticFields
from .core.grid import GridSpec, YeeGrid
from .core.simulation import Simulation
from .core.solver import FDTDSolver, MaxwellUpdater
from .geometry import Box, CustomShape, Cylinder, GeometryGroup, Material, Polygon, Shape, Sphere
from .materials.ade import ADEManager, ADESolver
from .materials.dispersion import (
    DebyeMaterial,
    DebyePole,
    DispersiveMaterial,
    DrudeMaterial,
    DrudePole,
    LorentzMaterial,
    LorentzPole,
    SellmeierMaterial,
)
from .monitors.dft import DFTMonitor
from .monitors.field import FieldMonitor
from .sources.base import Source
from .modes import ModeSolver, WaveguideMode
from .sources.gaussian import GaussianBeamSource
from .sources.mode import ModeSource
from .sources.plane_wave import PlaneWaveSource
from .sources.point import ElectricDipole, MagneticDipole, PointSource

__version__ = "0.1.0"

__all__ = [
    "backends", "boundaries", "core", "geometry", "materials", "monitors", "sources",
    "Simulation", "YeeGrid", "GridSpec", "FDTDSolver", "MaxwellUpdater",
    "ElectromagneticFields", "Backend", "get_backend", "set_backend",
    "list_available_backends", "Source", "PointSource", "ElectricDipole",
    "MagneticDipole", "GaussianBeamSource", "PlaneWaveSource", "FieldMonitor",
    "DFTMonitor", "CPML", "PMLParams", "DispersiveMaterial", "LorentzMaterial",
    "DrudeMaterial", "DebyeMaterial", "SellmeierMaterial", "LorentzPole", "DrudePole",
    "DebyePole", "ADESolver", "ADEManager", "Shape", "Material", "Box", "Sphere",
    "Cylinder", "Polygon", "CustomShape", "GeometryGroup", "ModeSolver", "WaveguideMode",
    "ModeSource",
]
