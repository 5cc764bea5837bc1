"""GPU: the 2-D TM one-pass step of one simulation (csrc/tm2d.hip) against
the unfused 2-D kernels and against O1.

Grids with CPML on the four sides (one whose strip boundary at x = 128
falls inside the x layer, so the recomputed Hy right of a strip takes its
CPML term), without CPML, FP64 and FP32 with FP32 or FP16 Cb; tiles of 1 to 32
rows; odd run lengths (the first step in place) and runs split into
pieces; a plane-wave line source and a spectral line monitor.
"""

import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
D = 20e-9
FREQS = [187e12, 193e12, 200e12]


def _sim(nxp, nyp, t, dtype, storage="auto", seed=5):
    import prismo_amd as pa

    rng = np.random.default_rng(seed)
    s = 1 - 1e-9
    g = pa.YeeGrid(pa.GridSpec(size=(nxp * D * s, nyp * D * s, 0.0), resolution=1 / D,
                               boundary_layers=t))
    nx, ny, _ = g.dimensions
    eps = np.full((nx, ny, 1), 2.085)
    yc = ny // 2
    eps[:, yc - 5:yc + 6] = 12.1
    eps *= rng.uniform(1.0, 1.05, size=eps.shape)
    kw = {"pml": pa.PMLParams(t)} if t else {}
    sim = pa.Simulation(size=(nxp * D * s, nyp * D * s, 0.0), resolution=1 / D, pml_layers=t,
                        courant_factor=0.99, scheme="yee", material_arrays={"eps_rel": eps},
                        dtype=dtype, coefficient_storage=storage, **kw)
    y0 = (yc - t) * D
    sim.add_source(pa.PlaneWaveSource(center=(6 * D, y0, 0), size=(0, 14 * D, 0),
                                      direction="+x", polarization="z", frequency=193e12,
                                      pulse_width=5e-15))
    # near enough for the pulse to reach it within the runs below
    mon = pa.FieldMonitor(center=(min(nxp - 6, 40) * D, y0, 0), size=(0, 16 * D, 0),
                          components=["Ez", "Hy"], frequencies=FREQS, time_domain=False)
    sim.add_monitor(mon)
    return sim, eps, mon


def _oracle(sim, t):
    """O1 fed the Cb the device applies (the engine's FP64 copy, FP16-rounded
    for FP16 storage), and O0's RegionDFT for the monitor."""
    from oracle import compat as O0
    from oracle.compat import RegionDFT
    from oracle.yee import YeeOracle

    eng = sim.solver.updater.engine_for(sim.fields)
    nx, ny, _ = sim.grid.dimensions
    eps = np.asarray(sim._material_arrays["eps_rel"])
    _, _, _, Db = O0.coefficients(eps, np.ones_like(eps), np.zeros_like(eps),
                                  np.zeros_like(eps), sim.dt)
    o1 = YeeOracle(nx, ny, 1, sim.grid.spacing, sim.dt, {"Ez": eng.cb_host["Ez"]},
                   float(Db.flat[0]), pml=(t, t, t, t, 0, 0) if t else None)
    mon = sim.monitors[0]
    ref = RegionDFT({c: mon._monitor_region[c] for c in mon.components}, mon.frequencies)
    return o1, ref


def _drive(o1, ref, sim, n):
    tt = 0.0
    for _ in range(n):
        o1.step()
        tt += sim.dt
        for s in sim.sources:
            s.update_fields(o1.F, tt, sim.dt)
        ref.update(o1.F, tt, sim.dt)


def _fields(sim):
    return {c: sim.fields.to_numpy(c) for c in ("Ez", "Hx", "Hy")}


def _run(shape, t, dtype, storage, splits, fused=True, rows=8):
    sim, eps, mon = _sim(*shape, t, dtype, storage)
    eng = sim.solver.updater.engine_for(sim.fields)
    assert eng.fused, "2-D TM yee plan should take the one-pass step"
    eng.ctx.set_tuning(1 if fused else 0, 4, rows, 0, 2)
    assert eng.ctx.fused_steps() == (1 if fused else 0)
    for n in splits:
        sim.run_steps(n)
    return sim, eps, mon


@pytest.mark.parametrize("shape,t", [((100, 60), 20), ((150, 40), 8), ((90, 70), 0)])
@pytest.mark.parametrize("rows", [1, 2, 3, 8, 32])
def test_tm2d_equals_unfused(mi355x, shape, t, rows):
    a, _, ma = _run(shape, t, np.float64, "auto", [131], fused=False)
    b, _, mb = _run(shape, t, np.float64, "auto", [131], fused=True, rows=rows)
    fa, fb = _fields(a), _fields(b)
    for c in fa:
        err = rel_l2(fb[c], fa[c])
        assert err <= 1e-13, (c, err)
    for c in ma.components:
        for f in FREQS:
            assert rel_l2(mb.get_frequency_data(c, f), ma.get_frequency_data(c, f)) <= 1e-13


@pytest.mark.parametrize("storage", ["auto", "fp16"])
@pytest.mark.parametrize("shape,t", [((100, 60), 20), ((90, 70), 0)])
def test_tm2d_fp32_equals_unfused(mi355x, shape, t, storage):
    """FP32 fields (FP32 or FP16 Cb): the one-pass kernel against the
    unfused pair; the same arithmetic per cell, but another instruction
    stream, whose multiply-adds the compiler may fuse differently."""
    a, _, _ = _run(shape, t, np.float32, storage, [131], fused=False)
    b, _, _ = _run(shape, t, np.float32, storage, [131], fused=True, rows=3)
    fa, fb = _fields(a), _fields(b)
    for c in fa:
        err = rel_l2(fb[c], fa[c])
        assert err <= 1e-6, (c, err)


@pytest.mark.parametrize("dtype,storage,bar", [(np.float64, "auto", 1e-10),
                                               (np.float32, "auto", 1e-5),
                                               (np.float32, "fp16", 1e-5)])
@pytest.mark.parametrize("shape,t,splits", [((100, 60), 20, [151]),
                                            ((150, 40), 8, [40, 33, 78]),
                                            ((90, 70), 0, [91])])
def test_tm2d_vs_oracle(mi355x, dtype, storage, bar, shape, t, splits):
    sim, _, mon = _run(shape, t, dtype, storage, splits)
    o1, ref = _oracle(sim, t)
    _drive(o1, ref, sim, sum(splits))
    got = _fields(sim)
    err = rel_l2(np.concatenate([got[c].ravel() for c in got]),
                 np.concatenate([o1.F[c].ravel() for c in got]))
    print(f"2-D one-pass {np.dtype(dtype).name} {shape} vs O1: {err:.2e}")
    assert err <= bar
    spec = np.stack([mon.get_frequency_data(c, f) for c in mon.components for f in FREQS])
    sref = np.stack([ref.data[c][f] for c in mon.components for f in FREQS])
    assert rel_l2(spec, sref) <= bar
