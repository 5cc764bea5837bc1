"""GPU: pm_configure_static, the configure entry point that allocates
nothing (the library's own per-device scratch in place of the caller's
plan.scratch), gives bit for bit what pm_configure gives, on two calls in a
row: a 3-D FP32 run with CPML, a Lorentz box, a point source and a spectral
monitor (the ADE, source and sample-box tables live in the scratch), then a
smaller 2-D reference-arithmetic run (its gating flags live there too)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
D = 20e-9
S = 1 - 1e-9


def _run_3d(static):
    import prismo_amd as pa

    n, t = 32, 5
    rng = np.random.default_rng(31)
    eps = rng.choice([1.0, 2.085, 12.1], size=(n, n, n))
    sim = pa.Simulation(size=((n - 2 * t) * D * S,) * 3, resolution=1 / D, pml_layers=t,
                        courant_factor=0.99, scheme="yee", pml=pa.PMLParams(t),
                        material_arrays={"eps_rel": eps}, dtype=np.float32)
    si = pa.LorentzMaterial(7.98, [pa.LorentzPole(omega_0=1.1e16, delta_epsilon=3.69,
                                                  gamma=1e13)])
    sim.add_dispersive_box(si, (4 * D, 4 * D, 4 * D), (14 * D, 12 * D, 10 * D))
    sim.add_source(pa.ElectricDipole(position=(8 * D, 8 * D, 8 * D), polarization="z",
                                     frequency=193e12, pulse_width=4e-15))
    mon = pa.FieldMonitor(center=(12 * D, 8 * D, 8 * D), size=(0.0, 8 * D, 6 * D),
                          components=["Ez", "Hy"], frequencies=[188e12, 196e12, 204e12],
                          time_domain=False)
    sim.add_monitor(mon)
    eng = sim.solver.updater.engine_for(sim.fields)
    if static:
        eng.use_static_scratch()
    sim.run_steps(61)
    return sim, mon


def _run_2d(static):
    import prismo_amd as pa

    sim = pa.Simulation(size=(24 * D * S, 16 * D * S, 0.0), resolution=1 / D, pml_layers=3,
                        courant_factor=0.99)
    sim.add_source(pa.PlaneWaveSource(center=(4 * D, 7 * D, 0.0), size=(0.0, 6 * D, 0.0),
                                      direction="+x", polarization="z", frequency=193e12,
                                      pulse_width=3e-15))
    mon = pa.FieldMonitor(center=(2 * D, 7 * D, 0.0), size=(0.0, 6 * D, 0.0),
                          components=["Ez", "Hx"], frequencies=[190e12, 199e12],
                          time_domain=True)
    sim.add_monitor(mon)
    eng = sim.solver.updater.engine_for(sim.fields)
    if static:
        eng.use_static_scratch()
    sim.run_steps(40)
    return sim, mon


def _same(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def _check(run):
    sim_s, mon_s = run(True)
    sim_c, mon_c = run(False)
    for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        a, b = sim_s.fields.to_numpy(c), sim_c.fields.to_numpy(c)
        assert _same(a, b), c
    assert np.max(np.abs(sim_c.fields.to_numpy("Ez"))) > 0
    for c in mon_c.components:
        for f in mon_c.frequencies:
            assert _same(mon_s.get_frequency_data(c, f), mon_c.get_frequency_data(c, f)), (c, f)
    return sim_s


def test_static_scratch_equals_caller_scratch(mi355x):
    s3 = _check(_run_3d)     # first call: the dispersive 3-D plan
    s2 = _check(_run_2d)     # second call, smaller: the 2-D plan with gating flags
    assert s2.grid.dimensions[0] * s2.grid.dimensions[1] < np.prod(s3.grid.dimensions)


def test_static_scratch_refuses_caller_scratch(mi355x):
    sim, _ = _run_2d(False)
    eng = sim.solver.updater.engine_for(sim.fields)
    assert eng.plan.scratch
    with pytest.raises(ValueError, match="scratch must be NULL"):
        eng.ctx.configure_static(eng.plan)
