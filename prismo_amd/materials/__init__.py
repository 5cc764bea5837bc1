from .ade import ADEManager, ADESolver
from .dispersion import (
    DebyeMaterial,
    DebyePole,
    DispersiveMaterial,
    DrudeMaterial,
    DrudePole,
    LorentzMaterial,
    LorentzPole,
    SellmeierMaterial,
    discrete_susceptibility,
)

__all__ = ["ADEManager", "ADESolver", "DispersiveMaterial", "LorentzMaterial",
           "DrudeMaterial", "DebyeMaterial", "SellmeierMaterial", "LorentzPole",
           "DrudePole", "DebyePole", "discrete_susceptibility"]
