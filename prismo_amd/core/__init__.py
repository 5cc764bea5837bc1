from .fields import ElectromagneticFields
from .grid import GridSpec, YeeGrid, grid_from_counts
from .simulation import Simulation
from .solver import FDTDSolver, MaxwellUpdater

__all__ = ["YeeGrid", "GridSpec", "ElectromagneticFields", "MaxwellUpdater", "FDTDSolver",
           "Simulation", "grid_from_counts"]
