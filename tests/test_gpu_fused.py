"""GPU: the fused ping-pong Yee step (3-D, one x-row per wave64) against
the O1 oracle, across tile shapes, precisions and dispersive regions."""

import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

SEED = 20261202
D = 20e-9


def _sim(nx, ny, nz, t, dtype=np.float64, storage="auto", seed=SEED, nsrc=1):
    import prismo_amd as pa

    rng = np.random.default_rng(seed)
    eps = rng.choice([1.0, 2.085, 12.1], size=(nx // 4 + 1, ny // 4 + 1, nz // 4 + 1))
    eps = np.repeat(np.repeat(np.repeat(eps, 4, 0), 4, 1), 4, 2)[:nx, :ny, :nz]
    s = 1 - 1e-9
    sim = pa.Simulation(size=((nx - 2 * t) * D * s, (ny - 2 * t) * D * s, (nz - 2 * t) * D * s),
                        resolution=1 / D, pml_layers=t, courant_factor=0.99, scheme="yee",
                        pml=pa.PMLParams(t), material_arrays={"eps_rel": eps}, dtype=dtype,
                        coefficient_storage=storage)
    sim.add_source(pa.ElectricDipole(position=(10 * D, 7 * D, 6 * D), polarization="z",
                                     frequency=193e12, pulse_width=4e-15))
    if nsrc > 1:
        sim.add_source(pa.ElectricDipole(position=(20 * D, 3 * D, 9 * D), polarization="x",
                                         frequency=200e12, pulse_width=4e-15, amplitude=0.7))
    return sim, eps


def _oracle(sim, eps, t, cb=None):
    from oracle import compat as O0
    from oracle.yee import YeeOracle

    nx, ny, nz = sim.grid.dimensions
    Ca, Cb, Da, Db = O0.coefficients(eps, np.ones_like(eps), np.zeros_like(eps),
                                     np.zeros_like(eps), sim.dt)
    if cb is None:
        cb = {"Ex": O0._ex(Cb), "Ey": O0._ey(Cb), "Ez": O0._ez(Cb)}
    return YeeOracle(nx, ny, nz, sim.grid.spacing, sim.dt, cb, float(Db.flat[0]), pml=(t,) * 6)


def _drive(o1, sim, n):
    tt = 0.0
    for _ in range(n):
        o1.step()
        tt += sim.dt
        for s in sim.sources:
            s.update_fields(o1.F, tt, sim.dt)


def _all(f):
    return np.concatenate([f[c].ravel() for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")])


def _fields(sim):
    return {c: sim.fields.to_numpy(c) for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")}


@pytest.mark.parametrize("shape,steps", [((100, 30, 28), 61), ((100, 30, 28), 60),
                                         ((256, 18, 20), 33), ((230, 20, 17), 40)])
def test_fused_fp64_matches_oracle(mi355x, shape, steps):
    sim, eps = _sim(*shape, t=5, nsrc=2)
    eng = sim.solver.updater.engine_for(sim.fields)
    assert eng.fused and eng.ctx.fused_active(), "shape should take the fused path"
    sim.run(steps * sim.dt * 0.9999)
    o1 = _oracle(sim, eps, 5)
    _drive(o1, sim, steps)
    got = _fields(sim)
    assert rel_l2(_all(got), _all(o1.F)) <= 1e-10
    for c in ("Ex", "Ey", "Ez"):
        assert rel_l2(got[c], o1.F[c]) <= 1e-9, c


def test_fused_tilings_agree(mi355x):
    """Tile shape and order never change the result beyond rounding: a tiling
    decides whether a cell is stepped by the interior instance of the kernel
    or the CPML-slab one, and the vector width changes the instruction
    stream, so the compiler may fuse different multiply-adds (<= 1e-13)."""
    ref = None
    for ty, kc, vec, order in ((4, 16, 1, 0), (2, 1, 1, 0), (4, 5, 1, 2), (2, 3, 2, 1),
                               (4, 7, 2, 0), (2, 16, 2, 2), (3, 16, 1, 2), (3, 4, 2, 0),
                               (3, 2, 1, 3)):
        sim, _ = _sim(128, 22, 19, t=4)
        eng = sim.solver.updater.engine_for(sim.fields)
        eng.ctx.set_tuning(True, ty, kc, vec, order)
        assert eng.ctx.fused_active()
        sim.run(24 * sim.dt * 0.9999)
        f = _all(_fields(sim))
        if ref is None:
            ref = f
        assert rel_l2(f, ref) <= 1e-13, (ty, kc, vec, order)


def test_fused_tilings_agree_wide_rows(mi355x):
    """256-cell rows: one lane per cell with 3 rows per workgroup is the
    largest workgroup (1024 lanes); it, the 2-row shape and two cells per
    lane agree to rounding (FP32 fields, FP16 Cb, ragged y and z)."""
    ref = None
    for ty, kc, vec, order in ((2, 16, 1, 2), (3, 16, 1, 2), (3, 5, 1, 0), (2, 8, 2, 1),
                               (3, 2, 1, 3)):
        sim, _ = _sim(256, 23, 21, t=4, dtype=np.float32, storage="fp16")
        eng = sim.solver.updater.engine_for(sim.fields)
        eng.ctx.set_tuning(True, ty, kc, vec, order)
        assert eng.ctx.fused_active(), (ty, kc, vec, order)
        sim.run(24 * sim.dt * 0.9999)
        f = _all(_fields(sim))
        if ref is None:
            ref = f
        assert rel_l2(f, ref) <= 1e-6, (ty, kc, vec, order)


def test_fused_equals_unfused(mi355x):
    out = []
    for fuse in (True, False):
        sim, _ = _sim(128, 22, 19, t=4)
        eng = sim.solver.updater.engine_for(sim.fields)
        eng.ctx.set_tuning(fuse, 4, 16)
        assert eng.ctx.fused_active() == fuse
        sim.run(30 * sim.dt * 0.9999)
        out.append(_all(_fields(sim)))
    assert rel_l2(out[0], out[1]) <= 1e-13


def test_fused_fp32_fp16_coefficients(mi355x):
    """FP32 fields with FP16 Cb storage vs the FP64 oracle fed the same
    dequantized coefficients: the error is arithmetic precision only."""
    sim, eps = _sim(128, 26, 24, t=5, dtype=np.float32, storage="fp16")
    eng = sim.solver.updater.engine_for(sim.fields)
    assert eng.fused and eng.coeff_kind == 2
    steps = 150
    sim.run(steps * sim.dt * 0.9999)
    o1 = _oracle(sim, eps, 5, cb=dict(eng.cb_host))
    _drive(o1, sim, steps)
    err = rel_l2(_all(_fields(sim)), _all(o1.F))
    assert err <= 1e-5, err


def test_fused_ade_lorentz_drude(mi355x):
    import prismo_amd as pa

    sim, eps = _sim(100, 24, 22, t=4)
    si = pa.LorentzMaterial(7.98, [pa.LorentzPole(omega_0=1.1e16, delta_epsilon=3.69, gamma=1e13)])
    au = pa.DrudeMaterial(1.0, 1.37e16, 1.07e14)
    boxes = sim.add_dispersive_box(si, (4 * D, 2 * D, 2 * D), (12 * D, 9 * D, 8 * D))
    boxes += sim.add_dispersive_box(au, (30 * D, 0.0, 0.0), (40 * D, 10 * D, 10 * D))
    eng = sim.solver.updater.engine_for(sim.fields)
    assert eng.fused and len(boxes) == 6
    steps = 90
    sim.run(steps * sim.dt * 0.9999)
    o1 = _oracle(sim, eps, 4)
    for reg in sim._ade:
        o1.add_ade(reg.comp, reg.box, reg.kind, reg.eps_inf, reg.c)
    _drive(o1, sim, steps)
    assert rel_l2(_all(_fields(sim)), _all(o1.F)) <= 1e-10


@pytest.mark.parametrize("ty,kc,vec,order", [(2, 8, 1, 2), (3, 5, 1, 3), (4, 16, 2, 2),
                                             (2, 3, 2, 3)])
def test_fused_ade_fp32_tilings(mi355x, ty, kc, vec, order):
    """FP32 fields with Lorentz and Drude boxes that cross tile rows and
    z-chunks, under several launch shapes, against the FP64 oracle fed the
    same coefficients: the difference is FP32 arithmetic only."""
    import prismo_amd as pa

    sim, eps = _sim(100, 24, 22, t=4, dtype=np.float32)
    si = pa.LorentzMaterial(7.98, [pa.LorentzPole(omega_0=1.1e16, delta_epsilon=3.69, gamma=1e13)])
    au = pa.DrudeMaterial(1.0, 1.37e16, 1.07e14)
    boxes = sim.add_dispersive_box(si, (4 * D, 2 * D, 2 * D), (12 * D, 9 * D, 8 * D))
    boxes += sim.add_dispersive_box(au, (30 * D, 0.0, 0.0), (40 * D, 10 * D, 10 * D))
    eng = sim.solver.updater.engine_for(sim.fields)
    eng.ctx.set_tuning(True, ty, kc, vec, order)
    assert eng.ctx.fused_active() and len(boxes) == 6
    steps = 60
    sim.run(steps * sim.dt * 0.9999)
    o1 = _oracle(sim, eps, 4)
    for reg in sim._ade:
        o1.add_ade(reg.comp, reg.box, reg.kind, reg.eps_inf, reg.c)
    _drive(o1, sim, steps)
    assert rel_l2(_all(_fields(sim)), _all(o1.F)) <= 1e-5


def test_autotune_restores_state(mi355x):
    from prismo_amd import autotune

    sim, _ = _sim(128, 22, 19, t=4, dtype=np.float32, storage="fp16")
    sim.run(10 * sim.dt * 0.9999)
    before = _all(_fields(sim))
    eng = sim.solver.updater.engine_for(sim.fields)
    res = autotune.tune(eng, steps=2, warmup=2, force=True,
                        candidates={"vec": (1, 2), "ty": (2, 4), "kchunk": (4,), "order": (2,)})
    assert res["fused"] and res["vec"] in (1, 2)
    after = _all(_fields(sim))
    assert np.array_equal(before.view(np.uint32), after.view(np.uint32))
    again = autotune.tune(eng)
    assert again.get("cached")
