"""GPU: the device path pinned by closed forms and hand-worked cases.

* Step order (src/prismo/core/simulation.py:147-164) of the product's
  Simulation on the device, both schemes: the 4 x 4-node case worked out by
  hand in tests/test_oracle_pins.py (sample 1 = g(t1), sample 2 =
  g(t1) (1 - m kappa) + g(t2), m = 1 for the reference stencil, 2 for Yee).
* The device spectra (csrc/dft_mfma.hip, the phasor table the engine
  builds) against the pure-tone closed form: samples x_n = A cos(w0 t_n)
  written into the sample matrix and summed with pm_dft_flush.
* DFTMonitor (src/prismo/monitors/dft.py:108-160): the region form and the
  reference's [:10, :10] placeholder corner against O0's RegionDFT over
  the same reference-arithmetic run, and the power spectrum.
* GaussianBeamSource (src/prismo/sources/gaussian.py:167-298): its device
  terms against its own host update driving O1.
* The strip workload (C1: 540 x 290 cells, 3000 steps, reference
  arithmetic) bit for bit against O0, spectra against RegionDFT.
"""

import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
D = 20e-9
SHRINK = 1 - 1e-9


# ---------------------------------------------------------------- step order
def _tiny_sim(scheme):
    import prismo_amd as pa
    from prismo_amd.sources import GaussianPulse

    sim = pa.Simulation(size=(2 * D * SHRINK, 2 * D * SHRINK, 0.0), resolution=1 / D,
                        pml_layers=1, courant_factor=0.99, scheme=scheme)
    assert sim.grid.dimensions == (4, 4, 1)
    src = pa.PointSource(position=(0.0, 0.0, 0.0), component="Ez",
                         waveform=GaussianPulse(frequency=193e12, pulse_width=2e-16))
    sim.add_source(src)
    mon = pa.FieldMonitor(center=(0.0, 0.0, 0.0), size=(0.0, 0.0, 0.0), components=["Ez"],
                          time_domain=True, frequencies=[190e12])
    sim.add_monitor(mon)
    assert src._source_box["Ez"] == ((1, 2), (1, 2), (0, 1))
    assert mon._monitor_box["Ez"] == ((1, 2), (1, 2), (0, 1))
    return sim, src, mon


@pytest.mark.parametrize("scheme,m", [("compat", 1), ("yee", 2)])
def test_step_order_two_steps_by_hand(mi355x, scheme, m):
    from oracle import compat as O0

    sim, src, mon = _tiny_sim(scheme)
    sim.run_steps(2)
    dt = sim.dt
    t1, t2 = dt, dt + dt
    _, cb, _, db = O0.coefficients(1.0, 1.0, 0.0, 0.0, dt)
    kappa = cb * db * (1 / D**2 + 1 / D**2)
    g1, g2 = src.waveform(t1), src.waveform(t2)
    x1 = g1
    x2 = g1 * (1 - m * kappa) + g2
    t, x = mon.get_time_data("Ez")
    assert list(t) == [t1, t2]
    np.testing.assert_allclose(x.ravel(), [x1, x2], rtol=1e-13, atol=0)
    f = 190e12
    want = (x1 * np.exp(-1j * 2 * np.pi * f * t1) + x2 * np.exp(-1j * 2 * np.pi * f * t2)) * dt
    got = complex(np.asarray(mon.get_frequency_data("Ez", f)).ravel()[0])
    assert abs(got - want) <= 1e-13 * abs(want)
    assert sim.current_time == t2 and sim.step_count == 2


# ------------------------------------------------------------ pure-tone DFT
def _tone_closed_form(amp, w0, w, dt, n):
    def geo(r):
        return r * (1 - r**n) / (1 - r)
    return amp * 0.5 * dt * (geo(np.exp(1j * (w0 - w) * dt)) + geo(np.exp(-1j * (w0 + w) * dt)))


@pytest.mark.parametrize("dtype,bar", [(np.float64, 1e-12), (np.float32, 1e-5)])
def test_device_dft_pure_tone_closed_form(mi355x, dtype, bar):
    import torch

    import prismo_amd as pa

    n, t = 16, 3
    sim = pa.Simulation(size=((n - 2 * t) * D * SHRINK,) * 3, resolution=1 / D, pml_layers=t,
                        courant_factor=0.99, scheme="yee", dtype=dtype)
    freqs = np.linspace(180e12, 210e12, 64)
    mon = pa.FieldMonitor(center=(0.0, 0.0, 0.0), size=(4 * D, 6 * D, 0.0),
                          components=["Ez", "Hy"], time_domain=False, frequencies=list(freqs))
    sim.add_monitor(mon)
    eng = sim.solver.updater.engine_for(sim.fields)
    N, T = 700, 128
    dt = sim.dt
    times = np.cumsum(np.full(N, dt))
    eng.configure_io([], [mon], times, step0=0, sample_T=T)
    ncols = eng.ncols
    assert ncols > 40
    rng = np.random.default_rng(7)
    amp = rng.standard_normal(ncols)
    w0 = 2 * np.pi * 193e12
    for b in range(0, N, T):
        rows = min(T, N - b)
        x = amp[None, :] * np.cos(w0 * (np.arange(b + 1, b + rows + 1) * dt))[:, None]
        eng.X[:rows].copy_(torch.from_numpy(x).to(eng.X.device, eng.X.dtype))
        eng.ctx.dft_flush(b, rows)
    eng.sync()
    spec = eng.spectra(eng.mon[0])
    order = sorted(eng.mon[0].cols.items(), key=lambda kv: kv[1][0])
    assert [c for c, _ in order] == ["Ez", "Hy"] and order[0][1][0] == 0
    got = np.concatenate([spec[c] for c, _ in order], axis=1)
    ref = np.stack([_tone_closed_form(amp, w0, 2 * np.pi * f, dt, N) for f in freqs])
    err = rel_l2(got, ref)
    print(f"pure tone {np.dtype(dtype).name}: rel L2 {err:.3e}")
    assert err <= bar


def test_dft_flush_rejects_bad_ranges(mi355x):
    import prismo_amd as pa
    from prismo_amd import _native as N

    sim = pa.Simulation(size=(10 * D * SHRINK, 10 * D * SHRINK, 0.0), resolution=1 / D,
                        pml_layers=2, courant_factor=0.99, scheme="yee")
    mon = pa.FieldMonitor(center=(0.0, 0.0, 0.0), size=(0.0, 4 * D, 0.0), components=["Ez"],
                          time_domain=False, frequencies=[190e12, 200e12])
    sim.add_monitor(mon)
    eng = sim.solver.updater.engine_for(sim.fields)
    times = np.cumsum(np.full(100, sim.dt))
    eng.configure_io([], [mon], times, step0=0, sample_T=32)
    with pytest.raises((ValueError, RuntimeError)):
        eng.ctx.dft_flush(20, 20)   # rows 20..39 run past the 32 rows of X
    with pytest.raises((ValueError, RuntimeError)):
        eng.ctx.dft_flush(96, 8)    # the phasor table ends at step 99
    eng.run(0, 10, N.RUN_SAMPLE | N.RUN_DFT)
    with pytest.raises((ValueError, RuntimeError)):
        eng.ctx.dft_flush(5, 10)    # overlaps the 10 samples the run left pending
    eng.ctx.dft_flush(0, 10)        # exactly the pending block: summed once
    eng.ctx.dft_flush(0, 0)         # nothing pending any more
    eng.sync()


# ----------------------------------------------------------------- DFTMonitor
def test_dft_monitor_region_and_placeholder(mi355x):
    import prismo_amd as pa
    from oracle import compat as O0

    sim = pa.Simulation(size=(30 * D * SHRINK, 20 * D * SHRINK, 0.0), resolution=1 / D,
                        pml_layers=4, courant_factor=0.99)
    # near the [:10, :10] corner, so the placeholder region sees the pulse
    src = pa.PlaneWaveSource(center=(2 * D, 4 * D, 0.0), size=(0.0, 8 * D, 0.0),
                             direction="+x", polarization="z", frequency=193e12,
                             pulse_width=3e-15)
    sim.add_source(src)
    freqs = [185e12, 193e12, 201e12]
    # the reference stencil carries the pulse towards -x only: both regions
    # sit at or below the source line
    reg = pa.DFTMonitor(center=(1 * D, 6 * D, 0.0), size=(6 * D, 8 * D, 0.0),
                        frequencies=freqs, components=["Ez", "Hy"])
    cor = pa.DFTMonitor(center=(12 * D, 9 * D, 0.0), size=(6 * D, 8 * D, 0.0),
                        frequencies=freqs, components=["Ez", "Hx"], compat_placeholder=True)
    sim.add_monitor(reg)
    sim.add_monitor(cor)
    steps = 90
    sim.run_steps(steps)
    nx, ny, nz = sim.grid.dimensions
    shp = (nx, ny, nz)
    C = dict(zip(("Ca", "Cb", "Da", "Db"),
                 O0.coefficients(np.ones(shp), np.ones(shp), np.zeros(shp), np.zeros(shp), sim.dt)))
    F = O0.zeros(nx, ny, nz, True)
    o_reg = O0.RegionDFT({c: reg._monitor_region[c] for c in reg.components}, freqs)
    # the reference's extraction: field[:10, :10] whatever the region (dft.py:157-160)
    o_cor = O0.RegionDFT({c: np.ix_(range(10), range(10)) for c in cor.components}, freqs)
    O0.run(F, C, sim.grid.spacing, True, sim.dt, steps,
           sources=[lambda F_, t: src.update_fields(F_, t, sim.dt)], monitors=[o_reg, o_cor])
    for c in F:
        assert np.array_equal(sim.fields.to_numpy(c), F[c]), c
    for mon, ref in ((reg, o_reg), (cor, o_cor)):
        assert mon._time_steps == steps
        for c in mon.components:
            got = np.stack([mon.get_frequency_data(c, i) for i in range(len(freqs))])
            want = np.stack([ref.data[c][f] for f in freqs])
            assert got.shape == want.shape, (c, got.shape, want.shape)
            assert rel_l2(got, want) <= 1e-12, c
            ps = np.array([np.sum(np.abs(ref.data[c][f]) ** 2) for f in freqs])
            assert ps.max() > 0, c
            np.testing.assert_allclose(mon.get_power_spectrum(c, normalize=False), ps, rtol=1e-11)
            np.testing.assert_allclose(mon.get_power_spectrum(c), ps / ps.max(), rtol=1e-11)
    assert reg.get_frequency_data("Ez", 0).shape == tuple(ix.size for ix in reg._monitor_region["Ez"])


# ----------------------------------------------------------- Gaussian beam
@pytest.mark.parametrize("dtype,bar", [(np.float64, 1e-10), (np.float32, 1e-5)])
def test_gaussian_beam_device_terms_vs_host_update(mi355x, dtype, bar):
    import prismo_amd as pa
    from tests.test_gpu_parity import _yee_oracle

    kw = dict(size=(40 * D * SHRINK, 30 * D * SHRINK, 0.0), resolution=1 / D, pml_layers=6,
              courant_factor=0.99)
    sim = pa.Simulation(**kw, scheme="yee", pml=pa.PMLParams(6), dtype=dtype)
    beam = pa.GaussianBeamSource(center=(6 * D, 14 * D, 0.0), size=(0.0, 16 * D, 0.0),
                                 direction="x", polarization="z", frequency=193e12,
                                 beam_waist=5 * D, pulse_width=4e-15)
    sim.add_source(beam)
    mon = pa.FieldMonitor(center=(30 * D, 14 * D, 0.0), size=(0.0, 16 * D, 0.0),
                          components=["Ez", "Hy"], frequencies=[190e12, 197e12],
                          time_domain=False)
    sim.add_monitor(mon)
    steps = 300
    sim.run_steps(steps)
    nx, ny, _ = sim.grid.dimensions
    o1 = _yee_oracle(sim, np.ones((nx, ny, 1)), 6)
    ref = O0_region(mon)
    o1.run(steps, sources=[lambda F_, t: beam.update_fields(F_, t, sim.dt)], monitors=[ref])
    got = np.concatenate([sim.fields.to_numpy(c).ravel() for c in o1.F])
    want = np.concatenate([o1.F[c].ravel() for c in o1.F])
    err = rel_l2(got, want)
    print(f"gaussian beam {np.dtype(dtype).name}: fields rel L2 {err:.3e}")
    assert err <= bar
    for c in mon.components:
        for f in mon.frequencies:
            assert rel_l2(mon.get_frequency_data(c, f), ref.data[c][f]) <= bar, (c, f)
    # the beam profile reached the device: Hy carries E / 377
    assert np.max(np.abs(sim.fields.to_numpy("Hy"))) > 0


def O0_region(mon):
    from oracle.compat import RegionDFT

    return RegionDFT({c: mon._monitor_region[c] for c in mon.components}, mon.frequencies)


# ------------------------------------------------------------ C1 full size
def _same_with_nonfinite(a, b):
    """Bitwise on finite values, same places and signs of +-Inf and NaN (the
    bits of a NaN differ between host and device arithmetic)."""
    fa, fb = np.isfinite(a), np.isfinite(b)
    if not np.array_equal(fa, fb) or not np.array_equal(np.isnan(a), np.isnan(b)):
        return False
    inf = np.isinf(a)
    if not np.array_equal(a[inf], b[inf]):
        return False
    return np.array_equal(a[fa].view(np.uint64), b[fb].view(np.uint64))


def test_strip_full_size_bitwise(mi355x):
    """C1 (540 x 290 cells, 3000 steps, reference arithmetic).  The
    reference's stencil is unstable (one cell off between its E and H
    differences) and this configuration overflows after step 1854.  At step
    1800 every field is O0's bit for bit and the 11-frequency spectra are
    within 1e-10 of RegionDFT; at step 3000 the finite values are still O0's
    bit for bit and the overflowed cells (Inf, NaN) are the same cells."""
    from oracle import compat as O0
    from prismo_amd import workloads as W

    sim, mon, steps = W.strip_2d(scheme="compat")
    assert sim.grid.dimensions == (540, 290, 1) and steps == 3000
    nx, ny, nz = sim.grid.dimensions
    shp = (nx, ny, nz)
    eps = W._strip_eps()["eps_rel"]
    C = dict(zip(("Ca", "Cb", "Da", "Db"),
                 O0.coefficients(eps.reshape(shp), np.ones(shp), np.zeros(shp), np.zeros(shp),
                                 sim.dt)))
    F = O0.zeros(nx, ny, nz, True)
    src = sim.sources[0]
    ref = O0_region(mon)
    first = 1800
    sim.run_steps(first)
    with np.errstate(over="ignore", invalid="ignore"):
        t = O0.run(F, C, sim.grid.spacing, True, sim.dt, first,
                   sources=[lambda F_, tt: src.update_fields(F_, tt, sim.dt)], monitors=[ref])
    assert t == sim.current_time
    for c in F:
        a, b = sim.fields.to_numpy(c), F[c]
        assert np.isfinite(b).all(), c
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), c
    for c in mon.components:
        for f in mon.frequencies:
            assert rel_l2(mon.get_frequency_data(c, f), ref.data[c][f]) <= 1e-10, (c, f)
    sim.run_steps(steps - first)
    with np.errstate(over="ignore", invalid="ignore"):
        t = O0.run(F, C, sim.grid.spacing, True, sim.dt, steps - first, t0=t,
                   sources=[lambda F_, tt: src.update_fields(F_, tt, sim.dt)])
    assert t == sim.current_time and sim.step_count == steps
    assert not np.isfinite(F["Ez"]).all()
    for c in F:
        assert _same_with_nonfinite(sim.fields.to_numpy(c), F[c]), c
