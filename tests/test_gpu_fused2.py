"""GPU: two Yee steps per launch (temporal blocking, csrc/fused2_kernel.h)
against the one-step fused kernel and against O1.

The two-step pass computes every cell with the same per-step arithmetic as
the one-step kernel, so over a whole run the fields agree to FP32 rounding
of the instruction order (<= 1e-6 relative L2), and both are within the
FP32 accuracy bar of 1e-5 against O1 fed the same (FP16-rounded)
coefficients.  Cases:
padded rows of 128 and 256 cells, CPML on all faces with interior and slab
tiles, z-chunks of several heights and both tile orders, run lengths that
leave one to three steps before the pairs of passes, two point sources (one
inside the kernel for the first step of each pass, one after it), a plane
source whose E and H lines are both added inside, and runs split at odd
step counts.
"""

import numpy as np
import pytest

from tests.conftest import rel_l2
from tests.test_gpu_fused import _all, _drive, _fields, _oracle, _sim

pytestmark = pytest.mark.gpu
D = 20e-9


def _run(shape, steps, fuse, kchunk=16, order=3, storage="fp16", splits=None, plane=False):
    import prismo_amd as pa

    sim, eps = _sim(*shape, t=5, dtype=np.float32, storage=storage, nsrc=2)
    if plane:
        sim.add_source(pa.PlaneWaveSource(center=(12 * D, 6 * D, 7 * D),
                                          size=(0.0, 8 * D, 6 * D), direction="+x",
                                          polarization="y", frequency=196e12,
                                          pulse_width=4e-15))
    eng = sim.solver.updater.engine_for(sim.fields)
    if fuse == 2:
        eng.ctx.set_tuning(2, 4, kchunk, 0, order)
        assert eng.ctx.fused_steps() == 2, "plan should take two steps per launch"
    else:
        eng.ctx.set_tuning(True, 4, 8, 0, 2)
        assert eng.ctx.fused_steps() == 1
    for n in (splits or [steps]):
        sim.run_steps(n)
    return sim, eps


@pytest.mark.parametrize("shape", [(100, 34, 40), (200, 30, 36)])
@pytest.mark.parametrize("steps", [40, 41, 42, 43])
def test_two_step_equals_one_step(mi355x, shape, steps):
    one, _ = _run(shape, steps, 1)
    two, _ = _run(shape, steps, 2)
    err = rel_l2(_all(_fields(two)), _all(_fields(one)))
    print(f"two-step vs one-step {shape} {steps} steps: {err:.2e}")
    assert err <= 1e-6


@pytest.mark.parametrize("kchunk,order", [(8, 2), (12, 3), (24, 0), (64, 3)])
def test_two_step_tilings(mi355x, kchunk, order):
    shape = (200, 30, 36)
    one, _ = _run(shape, 44, 1)
    two, _ = _run(shape, 44, 2, kchunk=kchunk, order=order)
    assert rel_l2(_all(_fields(two)), _all(_fields(one))) <= 1e-6


def test_two_step_plane_source_and_splits(mi355x):
    """A plane source (E and H lines) inside the passes, and a run split into
    pieces of 5, 9, 4 and 22 steps (each piece its own head + passes)."""
    shape = (100, 34, 40)
    one, _ = _run(shape, 40, 1, plane=True)
    two, _ = _run(shape, 40, 2, plane=True, splits=[5, 9, 4, 22])
    assert two.step_count == 40
    assert rel_l2(_all(_fields(two)), _all(_fields(one))) <= 1e-6


@pytest.mark.parametrize("storage", ["fp16", "auto"])
def test_two_step_vs_oracle(mi355x, storage):
    """FP32 fields (Cb FP16 or FP32) vs the FP64 oracle fed the same
    coefficients: <= 1e-5 relative L2 over the six components."""
    from oracle import compat as O0

    shape, steps = (100, 34, 40), 120
    two, eps = _run(shape, steps, 2, storage=storage)
    eng = two.solver.updater.engine_for(two.fields)
    cb = None
    if storage == "fp16":
        _, Cb, _, _ = O0.coefficients(eps, np.ones_like(eps), np.zeros_like(eps),
                                      np.zeros_like(eps), two.dt)
        scale = eng.cb_scale
        cb = {c: (f(Cb) / scale).astype(np.float16).astype(np.float64) * scale
              for c, f in (("Ex", O0._ex), ("Ey", O0._ey), ("Ez", O0._ez))}
    o1 = _oracle(two, eps, 5, cb=cb)
    _drive(o1, two, steps)
    err = rel_l2(_all(_fields(two)), _all(o1.F))
    print(f"two-step {storage} vs oracle: {err:.2e}")
    assert err <= 1e-5


def test_two_step_not_taken_for_dispersive_or_monitors(mi355x):
    import prismo_amd as pa

    sim, _ = _sim(100, 34, 40, t=5, dtype=np.float32, storage="fp16")
    si = pa.LorentzMaterial(7.98, [pa.LorentzPole(omega_0=1.1e16, delta_epsilon=3.69,
                                                  gamma=1e13)])
    sim.add_dispersive_box(si, (2 * D, 2 * D, 2 * D), (12 * D, 10 * D, 10 * D))
    eng = sim.solver.updater.engine_for(sim.fields)
    eng.ctx.set_tuning(2, 4, 16, 0, 3)
    assert eng.ctx.fused_steps() == 1
    # a monitored run keeps the one-step kernel and still matches
    res = []
    for fuse in (1, 2):
        s2, _ = _sim(100, 34, 40, t=5, dtype=np.float32, storage="fp16")
        mon = pa.FieldMonitor(center=(10 * D, 8 * D, 8 * D), size=(0.0, 6 * D, 6 * D),
                              components=["Ez"], frequencies=[190e12, 200e12],
                              time_domain=False)
        s2.add_monitor(mon)
        e2 = s2.solver.updater.engine_for(s2.fields)
        e2.ctx.set_tuning(fuse, 4, 16, 0, 3)
        s2.run_steps(30)
        res.append((_all(_fields(s2)), mon.get_frequency_data("Ez", 190e12)))
    assert rel_l2(res[0][0], res[1][0]) <= 1e-6
    assert rel_l2(res[0][1], res[1][1]) <= 1e-6


def test_autotune_two_step_restores_state(mi355x):
    from prismo_amd import autotune

    sim, _ = _sim(200, 30, 36, t=5, dtype=np.float32, storage="fp16")
    sim.run_steps(6)
    eng = sim.solver.updater.engine_for(sim.fields)
    before = _all(_fields(sim)).copy()
    psi = [t.clone() for t in eng._psi_tensors()]
    res = autotune.tune(eng, steps=4, warmup=4, force=True,
                        candidates={"vec": (2,), "ty": (2,), "kchunk": (16,), "order": (3,),
                                    "two_step": {"kchunk": (16, 32), "order": (3,)}})
    assert res["fused"] and res["steps"] in (1, 2)
    assert np.array_equal(_all(_fields(sim)), before)
    for a, b in zip(eng._psi_tensors(), psi):
        assert bool((a == b).all())
