"""GPU: pm_run with PM_RUN_GRAPH (blocks of steps replayed as captured HIP
graphs with a device step clock) against the same runs launched step by
step: bitwise equal fields, spectra and time-domain samples, and a graph
run against O1.  (Small grids take graph replays by default, so the 2-D
parity tests against O0 and O1 run through them too.)

Cases: the 2-D strip in the reference arithmetic (compat, FP64; also with
every block starting on the second of its two gate flag sets) and in the
textbook scheme with CPML; a small 3-D grid with CPML in FP32 (fused
ping-pong kernel) with two dipoles, a plane-wave source and a spectral
monitor; odd run lengths (the in-place first step), runs split at steps that
are not block starts, and a time-domain monitor whose rows come back per
block.
"""

import numpy as np
import pytest

from tests.conftest import rel_l2
from tests.test_gpu_fused import D, _all, _fields, _sim

pytestmark = pytest.mark.gpu


def _strip(scheme, steps):
    from prismo_amd.workloads import strip_2d

    sim, mon, _ = strip_2d(scheme=scheme, dtype=np.float64, steps=steps)
    return sim, mon


def _spec(mon):
    return np.stack([mon.get_frequency_data(c, f) for c in mon.components
                     for f in mon.frequencies])


@pytest.mark.parametrize("scheme", ["compat", "yee"])
def test_strip_graph_equals_direct(mi355x, scheme):
    out = []
    for graph in (False, True):
        sim, mon = _strip(scheme, 700)
        sim.graph = graph
        eng = sim.solver.updater.engine_for(sim.fields)
        eng.ctx.set_timing(True)
        for n in (301, 399):
            sim.run_steps(n)
        g_ms, g_n = eng.ctx.timing(6)
        eng.ctx.set_timing(False)
        print(f"{scheme} graph={graph}: {g_n} replays, {g_ms:.4f} ms each")
        assert (g_n > 0) == graph
        out.append((_fields(sim), _spec(mon), sim.current_time))
    (fa, sa, ta), (fb, sb, tb) = out
    assert ta == tb
    for c in ("Ez", "Hx", "Hy"):
        assert np.array_equal(fa[c], fb[c]), c
    assert np.array_equal(sa, sb)


def test_compat_strip_graph_blocks_on_odd_flag_set(mi355x):
    """2-D compat steps alternate between two sets of gate flags; an extra
    H half-step before the run makes every graph block start on the second
    set (the graphs are kept per starting set).  Graph replay equals direct
    launches bitwise."""
    out = []
    for graph in (False, True):
        sim, mon = _strip("compat", 700)
        sim.graph = graph
        sim.run_steps(40)
        up = sim.solver.updater
        up.update_magnetic_fields(sim.fields)   # one H without its E
        for n in (300, 173):
            sim.run_steps(n)
        out.append((_fields(sim), _spec(mon)))
    (fa, sa), (fb, sb) = out
    for c in ("Ez", "Hx", "Hy"):
        assert np.array_equal(fa[c], fb[c]), c
    assert np.array_equal(sa, sb)


def _small3d(graph, storage="fp16", plane=True, td=False):
    import prismo_amd as pa

    sim, _ = _sim(100, 30, 28, t=5, dtype=np.float32, storage=storage, nsrc=2)
    if plane:
        sim.add_source(pa.PlaneWaveSource(center=(12 * D, 6 * D, 7 * D),
                                          size=(0.0, 8 * D, 6 * D), direction="+x",
                                          polarization="y", frequency=196e12,
                                          pulse_width=4e-15))
    mon = pa.FieldMonitor(center=(30 * D, 8 * D, 8 * D), size=(0.0, 6 * D, 6 * D),
                          components=["Ez", "Hy"], frequencies=[190e12, 200e12],
                          time_domain=td)
    sim.add_monitor(mon)
    sim.graph = graph
    return sim, mon


@pytest.mark.parametrize("splits", [[41], [300], [37, 93, 170], [128, 129]])
def test_small3d_graph_equals_direct(mi355x, splits):
    out = []
    for graph in (False, True):
        sim, mon = _small3d(graph)
        eng = sim.solver.updater.engine_for(sim.fields)
        assert eng.ctx.fused_steps() == 1
        for n in splits:
            sim.run_steps(n)
        out.append((_all(_fields(sim)), _spec(mon)))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_time_domain_rows_graph_equals_direct(mi355x):
    out = []
    for graph in (False, True):
        sim, mon = _small3d(graph, plane=False, td=True)
        sim.run_steps(260)
        t, ez = mon.get_time_data("Ez")
        out.append((np.asarray(t), np.asarray(ez), _spec(mon)))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])


def test_graph_run_vs_oracle(mi355x):
    """FP32/FP16 3-D graph run against O1 fed the same FP16 Cb: <= 1e-5."""
    from oracle import compat as O0
    from tests.test_gpu_fused import _drive, _oracle

    sim, eps = _sim(100, 30, 28, t=5, dtype=np.float32, storage="fp16", nsrc=2)
    sim.graph = True
    steps = 270
    sim.run_steps(steps)
    eng = sim.solver.updater.engine_for(sim.fields)
    _, Cb, _, _ = O0.coefficients(eps, np.ones_like(eps), np.zeros_like(eps),
                                  np.zeros_like(eps), sim.dt)
    scale = eng.cb_scale
    cb = {c: (f(Cb) / scale).astype(np.float16).astype(np.float64) * scale
          for c, f in (("Ex", O0._ex), ("Ey", O0._ey), ("Ez", O0._ez))}
    o1 = _oracle(sim, eps, 5, cb=cb)
    _drive(o1, sim, steps)
    err = rel_l2(_all(_fields(sim)), _all(o1.F))
    print(f"graph run vs O1: {err:.2e}")
    assert err <= 1e-5
