"""CPU: pins for the O0 oracle (oracle/compat.py), the restatement of the
reference's time step (src/prismo/core/solver.py:113-556).

What pins it, each chosen so a plausible slip fails:
  * an element-by-element loop restatement of solver.py:167-456 written from
    the reference text with explicit indices (not slices): catches a wrong
    trim, an axis swap or a one-cell shift, bitwise;
  * hand-computed one-step responses to a single nonzero Ez (2-D and 3-D):
    the forward differences, the + sign of the 2-D Hx update (solver.py:338)
    and the trimmed ranges show up at known cells;
  * the 2-D gates (solver.py:321,367) incl. NaN semantics of max(abs());
  * linearity, zero-in/zero-out, the vacuum coefficients of
    tests/test_core_solver.py:140-176 and the grid shapes of the golden
    vectors.
"""

from pathlib import Path

import numpy as np
import pytest

from oracle import compat as O0

GOLD = np.load(Path(__file__).parent / "golden" / "reference_vectors.npz")


# ---------------------------------------------------------------- loop form
def _loops_h3d(F, Da, Db, d):
    Ex, Ey, Ez, Hx, Hy, Hz = (F[c].copy() for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"))
    nx, ny, nz = Da.shape
    dx, dy, dz = d
    out = {c: F[c].copy() for c in F}
    # ranges: the smallest of the H shape and the two curl shapes
    # (solver.py:189-190, 214-215, 238-239)
    for i in range(nx - 1):
        for j in range(ny - 2):
            for k in range(nz - 2):
                da = 0.5 * (Da[i, j, k] + Da[i + 1, j, k])
                db = 0.5 * (Db[i, j, k] + Db[i + 1, j, k])
                a = (Ez[i, j + 1, k] - Ez[i, j, k]) / dy
                b = (Ey[i, j, k + 1] - Ey[i, j, k]) / dz
                out["Hx"][i, j, k] = da * Hx[i, j, k] - db * (a - b)
    for i in range(nx - 2):
        for j in range(ny - 1):
            for k in range(nz - 2):
                da = 0.5 * (Da[i, j, k] + Da[i, j + 1, k])
                db = 0.5 * (Db[i, j, k] + Db[i, j + 1, k])
                a = (Ex[i, j, k + 1] - Ex[i, j, k]) / dz
                b = (Ez[i + 1, j, k] - Ez[i, j, k]) / dx
                out["Hy"][i, j, k] = da * Hy[i, j, k] - db * (a - b)
    for i in range(nx - 2):
        for j in range(ny - 2):
            for k in range(nz - 1):
                da = 0.5 * (Da[i, j, k] + Da[i, j, k + 1])
                db = 0.5 * (Db[i, j, k] + Db[i, j, k + 1])
                a = (Ey[i + 1, j, k] - Ey[i, j, k]) / dx
                b = (Ex[i, j + 1, k] - Ex[i, j, k]) / dy
                out["Hz"][i, j, k] = da * Hz[i, j, k] - db * (a - b)
    return out


def _avg4(A, i0, i1, j0, j1):
    return 0.25 * (A[i0] + A[i1] + A[j0] + A[j1])


def _loops_e3d(F, Ca, Cb, d):
    Ex, Ey, Ez, Hx, Hy, Hz = (F[c].copy() for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"))
    dx, dy, dz = d
    out = {c: F[c].copy() for c in F}
    for (i, j, k) in np.ndindex(*Ex.shape):
        idx = ((i, j, k), (i, j + 1, k), (i, j, k + 1), (i, j + 1, k + 1))
        ca = _avg4(Ca, *idx)
        cb = _avg4(Cb, *idx)
        a = (Hz[i, j + 1, k] - Hz[i, j, k]) / dy
        b = (Hy[i, j, k + 1] - Hy[i, j, k]) / dz
        out["Ex"][i, j, k] = ca * Ex[i, j, k] + cb * (a - b)
    for (i, j, k) in np.ndindex(*Ey.shape):
        idx = ((i, j, k), (i + 1, j, k), (i, j, k + 1), (i + 1, j, k + 1))
        ca = _avg4(Ca, *idx)
        cb = _avg4(Cb, *idx)
        a = (Hx[i, j, k + 1] - Hx[i, j, k]) / dz
        b = (Hz[i + 1, j, k] - Hz[i, j, k]) / dx
        out["Ey"][i, j, k] = ca * Ey[i, j, k] + cb * (a - b)
    for (i, j, k) in np.ndindex(*Ez.shape):
        idx = ((i, j, k), (i + 1, j, k), (i, j + 1, k), (i + 1, j + 1, k))
        ca = _avg4(Ca, *idx)
        cb = _avg4(Cb, *idx)
        a = (Hy[i + 1, j, k] - Hy[i, j, k]) / dx
        b = (Hx[i, j + 1, k] - Hx[i, j, k]) / dy
        out["Ez"][i, j, k] = ca * Ez[i, j, k] + cb * (a - b)
    return out


def _loops_h2d(F, Da, Db, d):
    dx, dy = d[0], d[1]
    out = {c: F[c].copy() for c in F}
    Ez, Hx, Hy, Hz, Ex, Ey = (F[c] for c in ("Ez", "Hx", "Hy", "Hz", "Ex", "Ey"))
    if np.max(np.abs(Ez)) > 0:
        for i in range(Hx.shape[0]):
            for j in range(Ez.shape[1] - 1):
                da = 0.5 * (Da[i, j] + Da[i + 1, j])
                db = 0.5 * (Db[i, j] + Db[i + 1, j])
                out["Hx"][i, j] = da * Hx[i, j] + db * ((Ez[i, j + 1] - Ez[i, j]) / dy)
        for i in range(Ez.shape[0] - 1):
            for j in range(Hy.shape[1]):
                da = 0.5 * (Da[i, j] + Da[i, j + 1])
                db = 0.5 * (Db[i, j] + Db[i, j + 1])
                out["Hy"][i, j] = da * Hy[i, j] - db * ((Ez[i + 1, j] - Ez[i, j]) / dx)
    if np.max(np.abs(Ex)) > 0 or np.max(np.abs(Ey)) > 0:
        nx, ny = Hz.shape
        for i in range(nx):
            for j in range(ny):
                cx = (Ey[i + 1, j] - Ey[i, j]) / dx if i < nx - 2 else 0.0
                cy = (Ex[i, j + 1] - Ex[i, j]) / dy if j < ny - 2 else 0.0
                out["Hz"][i, j] = Da[i, j] * Hz[i, j] - Db[i, j] * (cx - cy)
    return out


def _loops_e2d(F, Ca, Cb, d):
    dx, dy = d[0], d[1]
    out = {c: F[c].copy() for c in F}
    Ez, Hx, Hy, Hz, Ex, Ey = (F[c] for c in ("Ez", "Hx", "Hy", "Hz", "Ex", "Ey"))
    for (i, j) in np.ndindex(*Ez.shape):
        ca = 0.25 * (Ca[i, j] + Ca[i + 1, j] + Ca[i, j + 1] + Ca[i + 1, j + 1])
        cb = 0.25 * (Cb[i, j] + Cb[i + 1, j] + Cb[i, j + 1] + Cb[i + 1, j + 1])
        out["Ez"][i, j] = ca * Ez[i, j] + cb * ((Hy[i + 1, j] - Hy[i, j]) / dx
                                                 - (Hx[i, j + 1] - Hx[i, j]) / dy)
    for (i, j) in np.ndindex(*Ex.shape):
        ca = 0.5 * (Ca[i, j] + Ca[i, j + 1])
        cb = 0.5 * (Cb[i, j] + Cb[i, j + 1])
        out["Ex"][i, j] = ca * Ex[i, j] + cb * ((Hz[i, j + 1] - Hz[i, j]) / dy)
    for (i, j) in np.ndindex(*Ey.shape):
        ca = 0.5 * (Ca[i, j] + Ca[i + 1, j])
        cb = 0.5 * (Cb[i, j] + Cb[i + 1, j])
        out["Ey"][i, j] = ca * Ey[i, j] - cb * ((Hz[i + 1, j] - Hz[i, j]) / dx)
    return out


def _rand_case(nx, ny, nz, is_2d, seed):
    rng = np.random.default_rng(seed)
    F = {c: rng.standard_normal(s) for c, s in O0.shapes(nx, ny, nz, is_2d).items()}
    shp = (nx, ny, 1) if is_2d else (nx, ny, nz)
    eps = rng.uniform(1.0, 12.0, shp)
    mu = rng.uniform(1.0, 2.0, shp)
    se = rng.uniform(0.0, 5e4, shp)
    sm = rng.uniform(0.0, 5e9, shp)
    dt = 1e-17
    Ca, Cb, Da, Db = O0.coefficients(eps, mu, se, sm, dt)
    # scale the curl coefficients so fields stay O(1) and every term counts
    C = {"Ca": Ca, "Cb": Cb * 1e-9 / Cb.mean(), "Da": Da, "Db": Db * 1e-9 / Db.mean()}
    return F, C


@pytest.mark.parametrize("shape", [(5, 4, 3), (3, 6, 4), (2, 2, 2)])
def test_o0_3d_equals_loop_restatement(shape):
    F, C = _rand_case(*shape, False, 1)
    d = (1e-8, 2e-8, 1.5e-8)
    want_h = _loops_h3d(F, C["Da"], C["Db"], d)
    O0.h3d(F, C, d)
    for c in F:
        assert np.array_equal(F[c], want_h[c]), c
    want_e = _loops_e3d(F, C["Ca"], C["Cb"], d)
    O0.e3d(F, C, d)
    for c in F:
        assert np.array_equal(F[c], want_e[c]), c


@pytest.mark.parametrize("shape", [(6, 5), (3, 7), (2, 3)])
@pytest.mark.parametrize("te", [False, True])
def test_o0_2d_equals_loop_restatement(shape, te):
    F, C = _rand_case(shape[0], shape[1], 1, True, 2)
    if not te:
        for c in ("Ex", "Ey", "Hz"):
            F[c][...] = 0
    d = (1e-8, 2e-8, 0.0)
    Da, Db = C["Da"][:, :, 0], C["Db"][:, :, 0]
    want_h = _loops_h2d(F, Da, Db, d)
    O0.h2d(F, C, d)
    for c in F:
        assert np.array_equal(F[c], want_h[c]), c
    want_e = _loops_e2d(F, C["Ca"][:, :, 0], C["Cb"][:, :, 0], d)
    O0.e2d(F, C, d)
    for c in F:
        assert np.array_equal(F[c], want_e[c]), c


# -------------------------------------------------------- hand-computed cases
def test_o0_2d_single_ez_hand_computed():
    """Ez = 1 at (2,2) only, Da = 1, Db = b: H picks up the forward
    differences, Hx with a + sign, Hy with a - sign; nothing else moves."""
    nx, ny = 6, 6
    F = O0.zeros(nx, ny, 1, True)
    F["Ez"][2, 2] = 1.0
    b, dx, dy = 0.5, 2.0, 4.0
    C = {"Ca": np.ones((nx, ny, 1)), "Cb": np.ones((nx, ny, 1)),
         "Da": np.ones((nx, ny, 1)), "Db": np.full((nx, ny, 1), b)}
    O0.h2d(F, C, (dx, dy, 0.0))
    want_hx = np.zeros_like(F["Hx"])
    want_hx[2, 1] = b * 1 / dy     # (Ez[2,2] - Ez[2,1]) / dy
    want_hx[2, 2] = b * -1 / dy    # (Ez[2,3] - Ez[2,2]) / dy
    want_hy = np.zeros_like(F["Hy"])
    want_hy[1, 2] = -b * 1 / dx    # -(Ez[2,2] - Ez[1,2]) / dx
    want_hy[2, 2] = -b * -1 / dx   # -(Ez[3,2] - Ez[2,2]) / dx
    assert np.array_equal(F["Hx"], want_hx)
    assert np.array_equal(F["Hy"], want_hy)
    assert not F["Hz"].any()


def test_o0_3d_single_ez_hand_computed():
    nx, ny, nz = 5, 5, 4
    F = O0.zeros(nx, ny, nz, False)
    F["Ez"][2, 2, 1] = 1.0
    b, d = 0.25, (1.0, 2.0, 4.0)
    C = {"Ca": np.ones((nx, ny, nz)), "Cb": np.ones((nx, ny, nz)),
         "Da": np.ones((nx, ny, nz)), "Db": np.full((nx, ny, nz), b)}
    O0.h3d(F, C, d)
    want_hx = np.zeros_like(F["Hx"])
    want_hx[2, 1, 1] = -b * (1 / d[1])
    want_hx[2, 2, 1] = -b * (-1 / d[1])
    want_hy = np.zeros_like(F["Hy"])
    want_hy[1, 2, 1] = -b * (-(1 / d[0]))
    want_hy[2, 2, 1] = -b * (-(-1 / d[0]))
    assert np.array_equal(F["Hx"], want_hx)
    assert np.array_equal(F["Hy"], want_hy)
    assert not F["Hz"].any()
    # a nonzero Ez in the last z plane is never seen by Hx/Hy (trimmed to nz-1)
    G = O0.zeros(nx, ny, nz, False)
    G["Ez"][2, 2, nz - 1] = 1.0
    O0.h3d(G, C, d)
    assert not G["Hx"].any() and not G["Hy"].any()


def test_o0_2d_gates_and_nan():
    nx, ny = 6, 5
    C = {"Ca": np.ones((nx, ny, 1)), "Cb": np.ones((nx, ny, 1)),
         "Da": np.ones((nx, ny, 1)), "Db": np.ones((nx, ny, 1))}
    F = O0.zeros(nx, ny, 1, True)
    F["Hx"][1, 1] = 3.0
    F["Hz"][2, 2] = 1.0
    before = {c: F[c].copy() for c in F}
    O0.h2d(F, C, (1.0, 1.0, 0.0))
    for c in F:  # both gates closed: Ez and Ex/Ey are all zero
        assert np.array_equal(F[c], before[c]), c
    F["Ez"][3, 2] = np.nan
    F["Ez"][1, 1] = 2.0
    O0.h2d(F, C, (1.0, 1.0, 0.0))
    assert np.array_equal(F["Hx"], before["Hx"])   # max(abs) is NaN: gate closed
    F["Ez"][3, 2] = 0.0
    O0.h2d(F, C, (1.0, 1.0, 0.0))
    assert not np.array_equal(F["Hx"], before["Hx"])


def test_o0_linear_and_zero_preserving():
    F, C = _rand_case(6, 5, 4, False, 3)
    G, _ = _rand_case(6, 5, 4, False, 4)
    a, b = 0.75, -1.5
    S = {c: a * F[c] + b * G[c] for c in F}
    for X in (F, G, S):
        for _ in range(3):
            O0.step(X, C, (1e-8, 1e-8, 1e-8), False)
    for c in F:
        ref = a * F[c] + b * G[c]
        assert np.max(np.abs(S[c] - ref)) <= 1e-12 * np.max(np.abs(ref)), c
    Z = O0.zeros(6, 5, 4, False)
    O0.step(Z, C, (1e-8, 1e-8, 1e-8), False)
    assert not any(Z[c].any() for c in Z)


def test_o0_vacuum_coefficients():
    dt = 3.8131497e-17
    one = np.ones((3, 3, 3))
    Ca, Cb, Da, Db = O0.coefficients(one, one, 0 * one, 0 * one, dt)
    assert np.all(Ca == 1.0) and np.all(Da == 1.0)
    assert np.all(Cb == dt / O0.EPS0) and np.all(Db == dt / O0.MU0)
    Ca, Cb, _, _ = O0.coefficients(11.9 * one, one, 0 * one, 0 * one, dt)
    assert np.allclose(Cb, dt / (O0.EPS0 * 11.9), rtol=1e-15)
    # lossy: E decays by Ca per step ~ exp(-sigma dt / eps)
    sig = 1e5
    Ca, _, _, _ = O0.coefficients(one, one, sig * one, 0 * one, dt)
    x = sig * dt / O0.EPS0
    assert abs(np.log(Ca[0, 0, 0]) + x) <= x**3


@pytest.mark.parametrize("tag", ["T3", "T2", "C2", "C1exact"])
def test_o0_shapes_are_the_reference_grid_shapes(tag):
    nx, ny, nz = (int(v) for v in GOLD[f"{tag}_dims"])
    is_2d = GOLD[f"{tag}_shape_Ez"].size == 2
    sh = O0.shapes(nx, ny, nz, is_2d)
    for c, s in sh.items():
        assert tuple(s) == tuple(GOLD[f"{tag}_shape_{c}"]), (tag, c)
