"""Auxiliary-differential-equation state for dispersive regions.

``ADESolver`` / ``ADEManager`` keep the reference's interface
(src/prismo/materials/ade.py:16-359): per-material P/J state updated from a
given E with the recursions of ``dispersion.py``.  In the reference they are
never coupled to the E update; in this package the coupled form runs inside
the E kernel (``Simulation.add_dispersive_box``), and these classes remain the
stand-alone state updaters, working on whatever arrays the backend holds.
"""

from __future__ import annotations

from ..backends.backend_manager import resolve_backend
from .dispersion import DebyeMaterial, DrudeMaterial, LorentzMaterial

EPS0 = 8.854187817e-12


class ADESolver:
    def __init__(self, material, dt: float, grid_shape: tuple, backend=None):
        self.material = material
        self.dt = dt
        self.grid_shape = grid_shape
        self.backend = resolve_backend(backend)
        self.eps0 = EPS0
        self.coeffs = material.get_ade_coefficients(dt)
        b = self.backend
        if isinstance(material, LorentzMaterial):
            self.P_current = [b.zeros(grid_shape, dtype=b.float64) for _ in self.coeffs["poles"]]
            self.P_previous = [b.zeros(grid_shape, dtype=b.float64) for _ in self.coeffs["poles"]]
        elif isinstance(material, DrudeMaterial):
            self.J_current = b.zeros(grid_shape, dtype=b.float64)
        elif isinstance(material, DebyeMaterial):
            self.P_current = b.zeros(grid_shape, dtype=b.float64)

    def update_polarization(self, E_field) -> None:
        m = self.material
        if isinstance(m, LorentzMaterial):
            for i, pc in enumerate(self.coeffs["poles"]):
                new = (pc["C0"] * E_field + pc["C1"] * E_field
                       + pc["C2"] * self.P_current[i] + pc["C3"] * self.P_previous[i])
                self.P_previous[i] = self.backend.copy(self.P_current[i])
                self.P_current[i] = new
        elif isinstance(m, DrudeMaterial):
            self.J_current = self.coeffs["C0"] * E_field + self.coeffs["C1"] * self.J_current
        elif isinstance(m, DebyeMaterial):
            self.P_current = self.coeffs["C0"] * E_field + self.coeffs["C1"] * self.P_current

    def get_polarization_current(self):
        m = self.material
        b = self.backend
        if isinstance(m, LorentzMaterial):
            tot = b.zeros(self.grid_shape, dtype=b.float64)
            for p in self.P_current:
                tot = tot + p
            return tot
        if isinstance(m, DrudeMaterial):
            return self.J_current
        if isinstance(m, DebyeMaterial):
            return self.P_current
        return b.zeros(self.grid_shape, dtype=b.float64)

    def get_total_permittivity(self, E_field):
        b = self.backend
        einf = self.coeffs.get("epsilon_inf", 1.0)
        P = self.get_polarization_current()
        mask = b.abs(E_field) > 1e-30
        eps = einf * b.ones(self.grid_shape, dtype=b.float64)
        if b.sum(mask) > 0:
            eps = b.where(mask, einf + P / (self.eps0 * E_field + 1e-30), eps)
        return eps


class ADEManager:
    def __init__(self, grid_shape: tuple, dt: float, backend=None):
        self.grid_shape = grid_shape
        self.dt = dt
        self.backend = resolve_backend(backend)
        self.ade_solvers: list = []

    def add_material_region(self, material, region_mask) -> None:
        s = ADESolver(material, self.dt, self.grid_shape, backend=self.backend)
        self.ade_solvers.append({"solver": s, "mask": self.backend.asarray(region_mask),
                                 "material": material})

    def update_all(self, E_field) -> None:
        for d in self.ade_solvers:
            d["solver"].update_polarization(E_field * d["mask"])

    def get_total_polarization(self):
        b = self.backend
        tot = b.zeros(self.grid_shape, dtype=b.float64)
        for d in self.ade_solvers:
            tot = tot + d["solver"].get_polarization_current() * d["mask"]
        return tot

    def get_effective_permittivity(self, E_field):
        b = self.backend
        eps = b.ones(self.grid_shape, dtype=b.float64)
        for d in self.ade_solvers:
            eps = b.where(d["mask"], d["solver"].get_total_permittivity(E_field * d["mask"]), eps)
        return eps
