"""FP64 update coefficients and their averages at each component's points.

``cell_coefficients`` is the formula of src/prismo/core/solver.py:113-133;
``component_average`` forms the averages of solver.py:458-533 (3-D 4-point
face averages for E, 2-point edge averages for H; 2-D variants incl. the
identity for Hz), with the same left-to-right summation, so precomputing them
once gives the bits the reference recomputes every step.
"""

from __future__ import annotations

import numpy as np

EPS0 = 8.854187817e-12
MU0 = 4 * np.pi * 1e-7


def cell_coefficients(eps_rel, mu_rel, sigma_e, sigma_m, dt):
    """(Ca, Cb, Da, Db) at cell nodes, FP64 (solver.py:119-130).  For
    lossless non-magnetic media the expressions reduce exactly (sigma = 0
    gives (1 - 0) / (1 + 0) = 1 and x / (1 + 0) = x in IEEE arithmetic), so
    that case skips the array passes that cannot change a bit, and Ca, Da
    and Db, constants there, are read-only views of one value."""
    eps_rel = np.asarray(eps_rel)
    if is_lossless_nonmagnetic(mu_rel, sigma_e, sigma_m):
        shape = np.broadcast_shapes(eps_rel.shape, np.shape(mu_rel), np.shape(sigma_e),
                                    np.shape(sigma_m))
        Cb = dt / (EPS0 * eps_rel)
        if Cb.shape != shape or Cb.dtype != np.float64:
            Cb = np.broadcast_to(Cb, shape).astype(np.float64)
        one = np.broadcast_to(np.float64(1.0), shape)
        Db = np.broadcast_to(np.float64(dt / (MU0 * 1.0)), shape)
        return one, Cb, one, Db
    eps = EPS0 * eps_rel
    sigma_dt_2eps = sigma_e * dt / (2 * eps)
    Ca = (1 - sigma_dt_2eps) / (1 + sigma_dt_2eps)
    Cb = (dt / eps) / (1 + sigma_dt_2eps)
    mu = MU0 * mu_rel
    sigma_m_dt_2mu = sigma_m * dt / (2 * mu)
    Da = (1 - sigma_m_dt_2mu) / (1 + sigma_m_dt_2mu)
    Db = (dt / mu) / (1 + sigma_m_dt_2mu)
    return Ca, Cb, Da, Db


def component_average(A: np.ndarray, comp: str, is_2d: bool) -> np.ndarray:
    """Average of a node array at the sample points of ``comp``.

    3-D: A is (nx, ny, nz); 2-D: A is (nx, ny) (the z slice of the 3-D array).
    """
    if is_2d:
        if comp == "Ex" or comp == "Hy":
            return 0.5 * (A[:, :-1] + A[:, 1:])
        if comp == "Ey" or comp == "Hx":
            return 0.5 * (A[:-1, :] + A[1:, :])
        if comp == "Ez":
            return 0.25 * (A[:-1, :-1] + A[1:, :-1] + A[:-1, 1:] + A[1:, 1:])
        if comp == "Hz":
            return A
        raise ValueError(comp)
    if comp == "Ex":
        return 0.25 * (A[:, :-1, :-1] + A[:, 1:, :-1] + A[:, :-1, 1:] + A[:, 1:, 1:])
    if comp == "Ey":
        return 0.25 * (A[:-1, :, :-1] + A[1:, :, :-1] + A[:-1, :, 1:] + A[1:, :, 1:])
    if comp == "Ez":
        return 0.25 * (A[:-1, :-1, :] + A[1:, :-1, :] + A[:-1, 1:, :] + A[1:, 1:, :])
    if comp == "Hx":
        return 0.5 * (A[:-1, :, :] + A[1:, :, :])
    if comp == "Hy":
        return 0.5 * (A[:, :-1, :] + A[:, 1:, :])
    if comp == "Hz":
        return 0.5 * (A[:, :, :-1] + A[:, :, 1:])
    raise ValueError(comp)


def _all_equal(a, v) -> bool:
    """np.all(a == v), reading one element of a constant view (all strides
    zero, as the defaults of unset material arrays are)."""
    a = np.asarray(a)
    if a.size and all(st == 0 for st in a.strides):
        return bool(a.flat[0] == v)
    return bool(np.all(a == v))


def is_lossless_nonmagnetic(mu_rel, sigma_e, sigma_m) -> bool:
    """True when Ca = Da = 1 exactly and Db is one constant."""
    return _all_equal(sigma_e, 0) and _all_equal(sigma_m, 0) and _all_equal(mu_rel, 1)
