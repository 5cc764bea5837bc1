"""CPU: closed-form pins of the running DFT and of the step order.

* RegionDFT (oracle/compat.py; src/prismo/monitors/field.py:138-143): a
  pure tone x_n = A cos(w0 t_n) sampled at t_n = n dt makes the running sum
  sum_n x_n exp(-i w t_n) dt two geometric series, summed here in closed
  form.  The sign of the exponent, the dt factor and the time base (the
  first sample at t = dt, after the first step) each change the result far
  beyond the tolerance, which the sensitivity checks below show.
* Step order (src/prismo/core/simulation.py:147-164): H update, E update,
  t += dt, sources at the new time, then monitors.  On a 4 x 4-node 2-D grid
  with one Ez dipole and a one-cell monitor the first two steps are worked
  out by hand: the sample of step 1 is g(t1) and that of step 2 is
  g(t1) (1 - m kappa) + g(t2), kappa = Cb Db (1/dx^2 + 1/dy^2), where the
  reference stencil (one cell off) gives m = 1 and the textbook Yee
  Laplacian m = 2.  Monitoring before the sources, sampling the time before
  the increment, or a missing H or E update all give other numbers.
"""

import numpy as np
import pytest

from oracle import compat as O0
from oracle.yee import YeeOracle

DT = 3.8131497e-17
F0 = 193e12


def _tone_closed_form(amp, w0, w, dt, n):
    """sum_{m=1..n} amp cos(w0 m dt) exp(-i w m dt) dt."""
    def geo(r):
        return n if r == 1 else r * (1 - r**n) / (1 - r)
    r1 = np.exp(1j * (w0 - w) * dt)
    r2 = np.exp(-1j * (w0 + w) * dt)
    return amp * 0.5 * dt * (geo(r1) + geo(r2))


def _tone_run(amp, w0, freqs, n, phase=0.0):
    F = {"Ez": np.zeros(amp.shape)}
    mon = O0.RegionDFT({"Ez": np.ix_(range(amp.shape[0]), range(amp.shape[1]))}, freqs)
    t = 0.0
    for _ in range(n):
        t += DT
        F["Ez"][...] = amp * np.cos(w0 * t + phase)
        mon.update(F, t, DT)
    return mon


@pytest.mark.parametrize("n", [1, 7, 1000])
def test_region_dft_pure_tone_closed_form(n):
    rng = np.random.default_rng(3)
    amp = rng.standard_normal((2, 3))
    w0 = 2 * np.pi * F0
    freqs = [180e12, 193e12, 205e12, F0 * (1 + 1e-3)]
    mon = _tone_run(amp, w0, freqs, n)
    for f in freqs:
        ref = _tone_closed_form(amp, w0, 2 * np.pi * f, DT, n)
        got = mon.data["Ez"][f]
        assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref)), (n, f)


def test_region_dft_pins_are_sensitive():
    """The closed form separates the plausible slips: exp(+i w t), a
    missing dt, and a time base one step early (t_n = (n - 1) dt)."""
    amp = np.array([[1.0, -0.5]])
    w0 = 2 * np.pi * F0
    n = 300
    f = 190e12
    w = 2 * np.pi * f
    ref = _tone_closed_form(amp, w0, w, DT, n)
    got = _tone_run(amp, w0, [f], n).data["Ez"][f]
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))
    scale = np.max(np.abs(ref))
    # conjugate phasor: the complex conjugate of the right sum (x is real)
    assert np.max(np.abs(np.conj(ref) - ref)) > 1e-3 * scale
    # no dt factor
    assert np.max(np.abs(ref / DT - ref)) > 1e-3 * scale
    # samples at t_n = (n - 1) dt: the same sum times exp(+i w dt), with the
    # tone delayed by one step
    early = _tone_closed_form(amp, w0, w, DT, n) * np.exp(1j * w * DT)
    early_tone = np.sum([amp * np.cos(w0 * m * DT) * np.exp(-1j * w * m * DT) * DT
                         for m in range(0, n)], axis=0)
    assert np.max(np.abs(early - ref)) > 1e-3 * scale
    assert np.max(np.abs(early_tone - ref)) > 1e-3 * scale


# ---------------------------------------------------------------- step order
D = 20e-9


def _g(t):
    """A Gaussian-pulse-shaped value per time (any distinct values serve)."""
    w = 4e-15
    return np.exp(-0.5 * ((t - 3 * w) / w) ** 2) * np.sin(2 * np.pi * F0 * t)


def _hand_samples(cb, db, m):
    t1 = DT
    t2 = DT + DT
    kappa = cb * db * (1 / D**2 + 1 / D**2)
    x1 = _g(t1)
    x2 = _g(t1) * (1 - m * kappa) + _g(t2)
    return (t1, t2), (x1, x2)


class _Rec:
    def __init__(self, cell, freq):
        self.cell = cell
        self.f = freq
        self.samples = []
        self.times = []
        self.dft = 0j

    def update(self, F, t, dt):
        x = float(F["Ez"][self.cell])
        self.samples.append(x)
        self.times.append(t)
        self.dft += x * np.exp(-1j * 2 * np.pi * self.f * t) * dt


def _coeffs():
    shp = (4, 4, 1)
    return dict(zip(("Ca", "Cb", "Da", "Db"),
                    O0.coefficients(np.ones(shp), np.ones(shp), np.zeros(shp), np.zeros(shp), DT)))


def _check(rec, cb, db, m):
    (t1, t2), (x1, x2) = _hand_samples(cb, db, m)
    assert rec.times == [t1, t2]
    np.testing.assert_allclose(rec.samples, [x1, x2], rtol=1e-13, atol=0)
    dft = (x1 * np.exp(-1j * 2 * np.pi * rec.f * t1) + x2 * np.exp(-1j * 2 * np.pi * rec.f * t2)) * DT
    assert abs(rec.dft - dft) <= 1e-13 * abs(dft)
    return x1, x2


def test_step_order_o0_two_steps_by_hand():
    C = _coeffs()
    F = O0.zeros(4, 4, 1, True)
    rec = _Rec((1, 1), 190e12)

    def src(F_, t):
        F_["Ez"][1, 1] += _g(t)

    t = O0.run(F, C, (D, D, 0.0), True, DT, 2, sources=[src], monitors=[rec])
    assert t == DT + DT
    cb = float(C["Cb"].flat[0])
    db = float(C["Db"].flat[0])
    x1, x2 = _check(rec, cb, db, m=1)
    # the other orders give other samples
    assert abs(x2 - (_g(DT) * (1 - 2 * cb * db * 2 / D**2) + _g(2 * DT))) > 1e-3 * abs(x2)


def test_step_order_o1_two_steps_by_hand():
    C = _coeffs()
    cb = float(C["Cb"].flat[0])
    db = float(C["Db"].flat[0])
    o1 = YeeOracle(4, 4, 1, (D, D, 0.0), DT, {"Ez": O0._ez2(C["Cb"][:, :, 0])}, db)
    rec = _Rec((1, 1), 200e12)

    def src(F_, t):
        F_["Ez"][1, 1] += _g(t)

    t = o1.run(2, sources=[src], monitors=[rec])
    assert t == DT + DT
    _check(rec, cb, db, m=2)
    # H after step 2 holds the curl of Ez(t1) only (H is updated before E)
    g1 = _g(DT)
    np.testing.assert_allclose(o1.F["Hx"][1, 1:3], [-db * g1 / D, db * g1 / D], rtol=1e-14)
    np.testing.assert_allclose(o1.F["Hy"][1:3, 1], [db * g1 / D, -db * g1 / D], rtol=1e-14)


def test_step_order_sensitivity():
    """Monitor before source, or sources at the time before the increment,
    change the samples of the hand-computed case."""
    C = _coeffs()
    cb = float(C["Cb"].flat[0])
    db = float(C["Db"].flat[0])
    (_, _), (x1, x2) = _hand_samples(cb, db, 1)
    # monitor first: samples (0, g1 (1 - kappa))
    assert abs(x1) > 0 and abs(x2 - _g(DT) * (1 - cb * db * 2 / D**2)) > 1e-3 * abs(x2)
    # source at t before the increment: g(0) = 0 and g(t1) at step 2
    assert abs(x1 - _g(0.0)) > 1e-3 * abs(x1)
