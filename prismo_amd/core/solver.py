"""MaxwellUpdater and FDTDSolver (src/prismo/core/solver.py:24-704).

Same constructors, attributes and methods.  Set-up stays FP64 on the host
(materials, Ca/Cb/Da/Db, the interpolation helpers); the half-steps run in
libpmfdtd on the 'mi355x' backend through ``StepEngine``.  Two schemes:

* ``scheme="compat"`` (default) reproduces the reference arithmetic bit for
  bit, including its one-cell-shifted H differences and 2-D gating;
* ``scheme="yee"`` is the textbook leapfrog with CPML (``pml=``) and coupled
  ADE regions, in FP64 or FP32 (field dtype), optionally FP16 coefficients.
"""

from __future__ import annotations

from typing import Optional

import numpy as np

from ..backends.backend_manager import resolve_backend
from .coefficients import EPS0, MU0, cell_coefficients
from .fields import ElectromagneticFields


def _as_cpml(grid, pml):
    from ..boundaries.pml import CPML, PMLParams

    if pml is None:
        return None
    if isinstance(pml, CPML):
        return pml
    if isinstance(pml, int):
        pml = PMLParams(thickness=pml)
    if isinstance(pml, PMLParams):
        return CPML(grid, pml)
    raise TypeError("pml must be None, an int thickness, PMLParams or CPML")


class MaxwellUpdater:
    def __init__(self, grid, dt: float, material_arrays: Optional[dict] = None, backend=None,
                 *, scheme: str = "compat", pml=None, ade=(),
                 coefficient_storage: str = "auto"):
        self.grid = grid
        self.dt = dt
        self.backend = resolve_backend(backend)
        self.eps0 = EPS0
        self.mu0 = MU0
        self.c = 299792458.0
        courant = self.grid.get_courant_number(dt)
        if courant >= 1.0:
            raise ValueError(
                f"Time step dt={dt:.2e} violates Courant condition (S={courant:.3f} >= 1)"
            )
        if scheme not in ("compat", "yee"):
            raise ValueError("scheme must be 'compat' or 'yee'")
        self.scheme = scheme
        self.cpml = _as_cpml(grid, pml)
        self.ade = list(ade)
        self.coefficient_storage = coefficient_storage
        self._initialize_materials(material_arrays)
        self._compute_update_coefficients()
        self._engines: dict = {}
        self.structures: list = []
        self.background_eps = 1.0
        self._has_material_arrays = bool(material_arrays)

    def set_structures(self, shapes, background_eps: float = 1.0) -> None:
        """Shapes painted on the device (later shapes on top of earlier ones,
        ``background_eps`` elsewhere) instead of the host material arrays;
        lossless, non-magnetic materials only."""
        if self._has_material_arrays:
            raise ValueError("give the structure either as material_arrays or as shapes")
        for sh in shapes:
            if getattr(sh.material, "mu_r", 1.0) != 1.0:
                raise NotImplementedError("device-painted shapes take non-magnetic materials")
        self.structures = list(shapes)
        self.background_eps = float(background_eps)
        self.invalidate()

    # -- set-up (host FP64) ---------------------------------------------------
    def _initialize_materials(self, material_arrays) -> None:
        shape = self.grid.dimensions
        ma = material_arrays or {}
        for name, default in (("eps_rel", 1.0), ("mu_rel", 1.0), ("sigma_e", 0.0),
                              ("sigma_m", 0.0)):
            if name in ma:
                a = np.asarray(ma[name], dtype=np.float64)
                if a.shape != shape:
                    a = np.broadcast_to(a, shape).astype(np.float64)
            else:
                # a constant: a read-only view of one value, no per-cell storage
                a = np.broadcast_to(np.float64(default), shape)
            setattr(self, name, a)

    def _compute_update_coefficients(self) -> None:
        self.Ca, self.Cb, self.Da, self.Db = cell_coefficients(
            self.eps_rel, self.mu_rel, self.sigma_e, self.sigma_m, self.dt
        )
        self.dx, self.dy, self.dz = self.grid.spacing

    # interpolation helpers with the reference's names (solver.py:458-533)
    def _interpolate_to_ex_points(self, a):
        return 0.25 * (a[:, :-1, :-1] + a[:, 1:, :-1] + a[:, :-1, 1:] + a[:, 1:, 1:])

    def _interpolate_to_ey_points(self, a):
        return 0.25 * (a[:-1, :, :-1] + a[1:, :, :-1] + a[:-1, :, 1:] + a[1:, :, 1:])

    def _interpolate_to_ez_points(self, a):
        return 0.25 * (a[:-1, :-1, :] + a[1:, :-1, :] + a[:-1, 1:, :] + a[1:, 1:, :])

    def _interpolate_to_hx_points(self, a):
        return 0.5 * (a[:-1, :, :] + a[1:, :, :])

    def _interpolate_to_hy_points(self, a):
        return 0.5 * (a[:, :-1, :] + a[:, 1:, :])

    def _interpolate_to_hz_points(self, a):
        return 0.5 * (a[:, :, :-1] + a[:, :, 1:])

    def _interpolate_to_ex_points_2d(self, a):
        return 0.5 * (a[:, :-1] + a[:, 1:])

    def _interpolate_to_ey_points_2d(self, a):
        return 0.5 * (a[:-1, :] + a[1:, :])

    def _interpolate_to_ez_points_2d(self, a):
        return 0.25 * (a[:-1, :-1] + a[1:, :-1] + a[:-1, 1:] + a[1:, 1:])

    def _interpolate_to_hz_points_2d(self, a):
        return a

    def _interpolate_to_hx_points_2d(self, a):
        return 0.5 * (a[:-1, :] + a[1:, :])

    def _interpolate_to_hy_points_2d(self, a):
        return 0.5 * (a[:, :-1] + a[:, 1:])

    # -- device stepping ------------------------------------------------------
    def engine_for(self, fields: ElectromagneticFields):
        from .engine import StepEngine

        key = id(fields)
        eng = self._engines.get(key)
        if eng is None or eng.fields is not fields:
            if self.structures:
                coeffs = ("device", self.structures, self.background_eps)
            else:
                coeffs = (self.Ca, self.Cb, self.Da, self.Db, self.mu_rel, self.sigma_e,
                          self.sigma_m)
            eng = StepEngine(
                self.grid, self.dt, fields, coeffs,
                scheme=self.scheme, cpml=self.cpml, ade=self.ade,
                coefficient_storage=self.coefficient_storage,
            )
            self._engines[key] = eng
        return eng

    def invalidate(self):
        self._engines.clear()

    def update_magnetic_fields(self, fields) -> None:
        self.engine_for(fields).update_h()

    def update_electric_fields(self, fields) -> None:
        self.engine_for(fields).update_e()

    def step(self, fields) -> None:
        eng = self.engine_for(fields)
        eng.update_h()
        eng.update_e()

    def get_time_step(self) -> float:
        return self.dt

    def get_courant_number(self) -> float:
        return self.grid.get_courant_number(self.dt)

    def __repr__(self) -> str:
        return (f"MaxwellUpdater(dt={self.dt:.2e}s, Courant={self.get_courant_number():.3f}, "
                f"grid={self.grid.dimensions})")


class FDTDSolver:
    def __init__(self, grid, dt: Optional[float] = None, material_arrays: Optional[dict] = None,
                 backend=None, *, fields: Optional[ElectromagneticFields] = None,
                 dtype=None, **updater_kw):
        self.grid = grid
        self.backend = resolve_backend(backend)
        if dt is None:
            dt = grid.suggest_time_step(safety_factor=0.95)
        self.fields = fields if fields is not None else ElectromagneticFields(
            grid, dtype=dtype, backend=self.backend)
        self.updater = MaxwellUpdater(grid, dt, material_arrays, backend=self.backend,
                                      **updater_kw)
        self.time = 0.0
        self.step_count = 0

    def run(self, total_time: float, callback=None) -> None:
        dt = self.updater.get_time_step()
        self.run_steps(int(np.ceil(total_time / dt)), callback)

    def run_steps(self, num_steps: int, callback=None) -> None:
        dt = self.updater.get_time_step()
        eng = self.updater.engine_for(self.fields)
        if callback is None:
            eng.run(self.step_count, num_steps, 0)
            for _ in range(num_steps):
                self.time += dt
                self.step_count += 1
            return
        for s in range(num_steps):
            eng.run(self.step_count, 1, 0)
            self.time += dt
            self.step_count += 1
            callback(self, s)

    def step(self, fields=None) -> None:
        if fields is None:
            fields = self.fields
        self.updater.step(fields)
        self.time += self.updater.get_time_step()
        self.step_count += 1

    def reset(self) -> None:
        self.fields.zero_fields()
        for eng in self.updater._engines.values():
            eng.reset_state()
        self.time = 0.0
        self.step_count = 0

    def get_simulation_info(self) -> dict:
        return {
            "time": self.time,
            "step_count": self.step_count,
            "dt": self.updater.get_time_step(),
            "courant_number": self.updater.get_courant_number(),
            "grid_dimensions": self.grid.dimensions,
            "is_2d": self.grid.is_2d,
            "field_energy": self.fields.get_field_energy(),
        }

    def __repr__(self) -> str:
        i = self.get_simulation_info()
        return (f"FDTDSolver(t={i['time']:.2e}s, steps={i['step_count']}, "
                f"E_total={i['field_energy']:.2e}J)")
