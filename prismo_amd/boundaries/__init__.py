from .pml import CPML, PMLParams

__all__ = ["CPML", "PMLParams"]
