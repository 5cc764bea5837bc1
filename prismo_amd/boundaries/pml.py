"""Convolutional PML: parameters and the 1-D recursion profiles.

``PMLParams`` keeps the reference's fields and defaults
(src/prismo/boundaries/pml.py:23-46) and adds ``reflection``, the design
reflection R used for the default sigma_max.  The reference computes graded
profiles (pml.py:117-151) and recursion coefficients (pml.py:184-257) but its
update is a no-op (pml.py:259-328); here the profiles drive the CPML terms
inside the stencil kernels.  Textbook CPML (Roden & Gedney 2000; Taflove &
Hagness, 3rd ed., sec. 7.9-7.11):

    rho in [0, 1] = depth into the layer (0 at the inner interface)
    sigma = sigma_max rho^m,  kappa = 1 + (kappa_max - 1) rho^m,
    alpha = alpha_max (1 - rho)^m
    sigma_max = -(m + 1) ln(R) / (2 eta0 d)           (d = layer thickness, m)
    b = exp(-(sigma / kappa + alpha) dt / eps0)
    c = sigma / (sigma kappa + kappa^2 alpha) (b - 1)
    psi <- b psi + c dF,   dF / kappa + psi replaces dF

The outer walls sit half a cell beyond the outermost H nodes (positions -1/2
and n - 1/2 in node units); a layer of t cells spans t cells inward from a
wall.  E-update terms are sampled at half nodes p = j + 1/2 and H-update terms
at nodes p = i, so rho(p) = (t - 1/2 - p) / t on the low side and
(p - (n - 1/2 - t)) / t on the high side.  Magnetic conductivity is matched
(sigma_m / mu0 = sigma / eps0), so both halves use the same formulas.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

EPS0 = 8.854187817e-12
MU0 = 4 * np.pi * 1e-7
ETA0 = float(np.sqrt(MU0 / EPS0))


@dataclass
class PMLParams:
    thickness: int
    sigma_max: Optional[float] = None
    kappa_max: float = 15.0
    alpha_max: float = 0.0
    polynomial_order: int = 3
    reflection: float = 1e-8


def sigma_max_for(params: PMLParams, d: float) -> float:
    if params.sigma_max is not None:
        return float(params.sigma_max)
    m = params.polynomial_order
    t = params.thickness
    return -(m + 1) * np.log(params.reflection) / (2.0 * ETA0 * t * d)


def depth(n: int, t_lo: int, t_hi: int, half: bool) -> np.ndarray:
    """rho at every index of an axis of n nodes (0 outside the layers)."""
    p = np.arange(n, dtype=np.float64) + (0.5 if half else 0.0)
    rho = np.zeros(n)
    if t_lo > 0:
        lo = (t_lo - 0.5 - p) / t_lo
        rho = np.where(lo > 0, lo, rho)
    if t_hi > 0:
        hi = (p - (n - 0.5 - t_hi)) / t_hi
        rho = np.where(hi > 0, hi, rho)
    return np.clip(rho, 0.0, 1.0)


_PROFILES: dict = {}


def axis_profiles(n: int, t_lo: int, t_hi: int, d: float, dt: float,
                  params: PMLParams) -> dict:
    """b, c, 1/kappa at half nodes (E terms) and nodes (H terms) for one axis.

    Indices outside both layers get b = 1, c = 0, 1/kappa = 1.  Each layer
    uses its own thickness for sigma_max and rho.  The profiles depend on
    the arguments only and are remembered per argument set (a sweep of many
    designs shares them), as read-only arrays.
    """
    key = (n, t_lo, t_hi, float(d), float(dt), params.sigma_max, params.kappa_max,
           params.alpha_max, params.polynomial_order, params.reflection)
    hit = _PROFILES.get(key)
    if hit is not None:
        return hit
    if len(_PROFILES) > 4096:
        _PROFILES.clear()
    out = _axis_profiles(n, t_lo, t_hi, d, dt, params)
    for a in out.values():
        a.setflags(write=False)
    _PROFILES[key] = out
    return out


def _axis_profiles(n: int, t_lo: int, t_hi: int, d: float, dt: float,
                   params: PMLParams) -> dict:
    m = params.polynomial_order
    out = {}
    for half, tag in ((True, "e"), (False, "h")):
        b = np.ones(n)
        c = np.zeros(n)
        ik = np.ones(n)
        for t, side in ((t_lo, "lo"), (t_hi, "hi")):
            if t <= 0:
                continue
            rho_all = depth(n, t if side == "lo" else 0, t if side == "hi" else 0, half)
            idx = np.arange(n)
            inside = (idx < t) if side == "lo" else (idx >= n - t)
            rho = rho_all[inside]
            smax = sigma_max_for(PMLParams(t, params.sigma_max, params.kappa_max,
                                           params.alpha_max, m, params.reflection), d)
            sigma = smax * rho**m
            kappa = 1.0 + (params.kappa_max - 1.0) * rho**m
            alpha = params.alpha_max * (1.0 - rho) ** m
            bb = np.exp(-(sigma / kappa + alpha) * dt / EPS0)
            denom = sigma * kappa + kappa**2 * alpha
            cc = np.where(denom > 0, sigma / np.where(denom > 0, denom, 1.0) * (bb - 1.0), 0.0)
            b[inside] = bb
            c[inside] = cc
            ik[inside] = 1.0 / kappa
        out["b" + tag] = b
        out["c" + tag] = c
        out["ik" + tag] = ik
    return out


def plane_wave_reflection(params: PMLParams, d: float) -> float:
    """Continuous-layer normal-incidence reflection exp(-2 eta0 int sigma dx)
    (kappa and alpha do not enter); equals ``reflection`` by construction
    when sigma_max takes its default."""
    t = params.thickness
    smax = sigma_max_for(params, d)
    integral = smax * t * d / (params.polynomial_order + 1)
    return float(np.exp(-2.0 * ETA0 * integral))


class CPML:
    """CPML description bound to a grid (the reference's constructor shape,
    src/prismo/boundaries/pml.py:67-93).  The time step arrives with
    ``profiles``: the layer is only meaningful for the dt the stencil uses."""

    def __init__(self, grid, params: PMLParams, backend=None):
        self.grid = grid
        self.params = params
        self.backend = backend
        t = params.thickness
        nx, ny, nz = grid.dimensions
        self.thickness = (t, t, t, t, 0 if grid.is_2d else t, 0 if grid.is_2d else t)
        for a, n in enumerate((nx, ny, nz)):
            if self.thickness[2 * a] + self.thickness[2 * a + 1] > n:
                raise ValueError(f"PML thickness {t} does not fit axis {a} of {n} nodes")

    def profiles(self, dt: float):
        g = self.grid
        out = []
        for a, (n, d) in enumerate(zip(g.dimensions, g.spacing)):
            tl, th = self.thickness[2 * a], self.thickness[2 * a + 1]
            if tl + th == 0:
                out.append(None)
            else:
                out.append(axis_profiles(n, tl, th, d, dt, self.params))
        return out

    def get_reflection_coefficient(self, angle: float = 0.0) -> float:
        d = max(s for s in self.grid.spacing if s > 0)
        r = plane_wave_reflection(self.params, d)
        return float(r ** np.cos(np.deg2rad(angle)))
