"""Mode solver (prismo_amd.modes) pinned by closed forms: the exact discrete
eigenvalues of a uniform closed window, and second-order convergence to the
analytic TE and TM slab-waveguide dispersion relations."""

import numpy as np
import pytest
from scipy.optimize import brentq

from prismo_amd.modes import ModeSolver

LAM = 1.55e-6
K0 = 2 * np.pi / LAM


@pytest.mark.parametrize("kind", ["TE", "TM"])
def test_uniform_window_exact_discrete_eigenvalue(kind):
    nx, ny, h = 60, 40, 20e-9
    ms = ModeSolver(LAM, np.arange(nx) * h, np.arange(ny) * h, np.full((nx, ny), 12.0))
    m = ms.solve(1, kind, guided_only=False)[0]
    b2 = (K0**2 * 12 - 4 / h**2 * np.sin(np.pi / (2 * (nx + 1)))**2
          - 4 / h**2 * np.sin(np.pi / (2 * (ny + 1)))**2)
    assert abs((m.neff.real * K0)**2 / b2 - 1) < 1e-12
    # fundamental: one sign over the window
    dom = m.Ex if kind == "TE" else m.Ey
    assert np.all(dom.real >= -1e-12) and abs(dom).max() == pytest.approx(1.0)


def _slab(h, kind):
    n1, n2, d, W = 3.48, 1.444, 220e-9, 2.4e-6
    ny = int(round(W / h))
    nc = int(round(d / h))
    j0 = (ny - nc) // 2
    eps = np.full((1, ny), n2**2)
    eps[0, j0:j0 + nc] = n1**2
    m = ModeSolver(LAM, np.array([0.0]), np.arange(ny) * h, eps).solve(1, kind)[0]
    # one node across x (spacing taken as 1): its Dirichlet term is -2
    got = np.sqrt((m.neff.real * K0)**2 + 2.0) / K0
    ratio = 1.0 if kind == "TE" else n1**2 / n2**2

    def f(b):
        kap = np.sqrt(K0**2 * n1**2 - b**2)
        gam = np.sqrt(b**2 - K0**2 * n2**2)
        return np.tan(kap * d / 2) - ratio * gam / kap

    lo = np.sqrt(max(K0**2 * n1**2 - (np.pi / d)**2 * 0.999, K0**2 * n2**2)) + 1
    want = brentq(f, lo, K0 * n1 * (1 - 1e-12)) / K0
    return got, want


@pytest.mark.parametrize("kind", ["TE", "TM"])
def test_slab_converges_to_analytic_dispersion(kind):
    e1 = [abs(g / w - 1) for g, w in (_slab(10e-9, kind), _slab(5e-9, kind))]
    assert e1[0] < 1e-3 and e1[1] < 3e-4
    assert 3.0 < e1[0] / e1[1] < 5.0          # second order
    # TE and TM differ by far more than the discretisation error
    assert abs(_slab(5e-9, "TE")[0] - _slab(5e-9, "TM")[0]) > 0.5


def test_strip_modes_ordered_and_carry_power():
    h = 20e-9
    ny, nz = 120, 80
    eps = np.ones((ny, nz))
    eps[:, :35] = 2.085
    eps[ny // 2 - 11:ny // 2 + 12, 35:46] = 11.7
    ms = ModeSolver(LAM, np.arange(ny) * h, np.arange(nz) * h, eps)
    te = ms.solve(2, "TE")
    assert len(te) >= 1 and te[0].neff.real > 2.0
    if len(te) > 1:
        assert te[0].neff.real > te[1].neff.real
    assert te[0].power > 0
    tm = ms.solve(1, "TM")[0]
    assert 1.444 < tm.neff.real < te[0].neff.real and tm.power > 0
    assert ms.get_mode(0) is tm
    with pytest.raises(IndexError):
        ms.get_mode(3)
    with pytest.raises(NotImplementedError):
        ms.solve(1, "vector")
