"""GPU: batched 2-D TM stepping (include/pmbatch.h, prismo_amd/core/batch.py)
against the O1 oracle, design by design.

Designs of uneven size (strip and chunk tails, several tiles per design)
with CPML, a PlaneWaveSource line (Ez and Hy) and a FieldMonitor port line
(Ez, Hy) with spectra.  FP64: fields and spectra within 1e-10 relative L2 of
O1; FP32 fields with FP16 Cb: within 1e-5 of O1 fed the same dequantised
Cb.  Graph replay of the step loop gives the same bits as eager launches.
"""

import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
D = 20e-9
# with 8-cell CPML: 86, 146, 61 and 68 cells along x; strips of 64 cells
# give tails, a one-strip design and (68) a strip whose right neighbour
# cell, x = 64, lies in the x CPML
SIZES = [(70, 40), (130, 25), (45, 90), (52, 33)]
FREQS = [187e12, 193e12, 200e12, 205e12]


def _design(nxp, nyp, t, dtype, seed, poison=False):
    import prismo_amd as pa

    rng = np.random.default_rng(seed)
    s = 1 - 1e-9
    g = pa.YeeGrid(pa.GridSpec(size=(nxp * D * s, nyp * D * s, 0.0), resolution=1 / D,
                               boundary_layers=t))
    nx, ny, _ = g.dimensions
    eps = np.full((nx, ny, 1), 2.085)
    yc = ny // 2 + int(rng.integers(-3, 4))
    eps[:, yc - 5:yc + 6] = 12.1
    eps *= rng.uniform(1.0, 1.05, size=eps.shape)
    if poison:
        eps[nx // 2, ny // 2, 0] = np.nan
    sim = pa.Simulation(size=(nxp * D * s, nyp * D * s, 0.0), resolution=1 / D, pml_layers=t,
                        courant_factor=0.99, scheme="yee", pml=pa.PMLParams(t),
                        material_arrays={"eps_rel": eps}, dtype=dtype)
    y0 = (yc - t) * D
    sim.add_source(pa.PlaneWaveSource(center=(6 * D, y0, 0), size=(0, 14 * D, 0), direction="+x",
                                      polarization="z", frequency=193e12, pulse_width=5e-15))
    mon = pa.FieldMonitor(center=((nxp - 6) * D, y0, 0), size=(0, 16 * D, 0),
                          components=["Ez", "Hy"], frequencies=FREQS, time_domain=False)
    sim.add_monitor(mon)
    return sim, eps


def _oracle(sim, eps, t, cb):
    from oracle import compat as O0
    from oracle.yee import YeeOracle
    from oracle.compat import RegionDFT

    nx, ny, _ = sim.grid.dimensions
    _, Cb, _, Db = O0.coefficients(eps, np.ones_like(eps), np.zeros_like(eps),
                                   np.zeros_like(eps), sim.dt)
    o1 = YeeOracle(nx, ny, 1, sim.grid.spacing, sim.dt, {"Ez": cb}, float(Db.flat[0]),
                   pml=(t, t, t, t, 0, 0))
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


def _run(dtype, storage, steps, graph, T=32, t=8, chunk=32):
    from prismo_amd.core.batch import BatchEngine

    built = [_design(nxp, nyp, t, dtype, 100 + k) for k, (nxp, nyp) in enumerate(SIZES)]
    sims = [b[0] for b in built]
    eng = BatchEngine(sims, dtype=dtype, coefficient_storage=storage, chunk=chunk, sample_T=T)
    eng.run_steps(steps, graph=graph)
    return eng, built


@pytest.mark.parametrize("steps,graph,chunk", [(150, False, 32), (151, False, 7), (150, True, 32)])
def test_batch_fp64_matches_oracle(mi355x, steps, graph, chunk):
    eng, built = _run(np.float64, "auto", steps, graph, chunk=chunk)
    for d, (sim, eps) in enumerate(built):
        o1, ref = _oracle(sim, eps, 8, eng.cb_host[d])
        _drive(o1, ref, sim, steps)
        got = eng.fields(d)
        for c in ("Ez", "Hx", "Hy"):
            assert rel_l2(got[c], o1.F[c]) <= 1e-10, (d, c)
        mon = sim.monitors[0]
        for c in mon.components:
            for f in mon.frequencies:
                assert rel_l2(mon.get_frequency_data(c, f), ref.data[c][f]) <= 1e-10, (d, c, f)


def test_batch_fp32_fp16_matches_oracle(mi355x):
    steps = 240
    eng, built = _run(np.float32, "fp16", steps, True)
    for d, (sim, eps) in enumerate(built):
        o1, ref = _oracle(sim, eps, 8, eng.cb_host[d])
        _drive(o1, ref, sim, steps)
        got = eng.fields(d)
        all_got = np.concatenate([got[c].ravel() for c in ("Ez", "Hx", "Hy")])
        all_ref = np.concatenate([o1.F[c].ravel() for c in ("Ez", "Hx", "Hy")])
        assert rel_l2(all_got, all_ref) <= 1e-5, d
        mon = sim.monitors[0]
        spec = np.stack([mon.get_frequency_data(c, f) for c in mon.components for f in FREQS])
        sref = np.stack([ref.data[c][f] for c in mon.components for f in FREQS])
        assert rel_l2(spec, sref) <= 1e-5, d


def test_batch_graph_equals_eager(mi355x):
    out = []
    for graph in (False, True):
        eng, built = _run(np.float32, "auto", 200, graph)
        f = [eng.fields(d) for d in range(len(built))]
        s = [np.stack([b[0].monitors[0].get_frequency_data(c, fr) for c in ("Ez", "Hy")
                       for fr in FREQS]) for b in built]
        out.append((f, s))
    (fa, sa), (fb, sb) = out
    for d in range(len(SIZES)):
        for c in ("Ez", "Hx", "Hy"):
            assert np.array_equal(fa[d][c], fb[d][c]), (d, c)
        assert np.array_equal(sa[d], sb[d]), d


def _spectra(built):
    return [np.stack([b[0].monitors[0].get_frequency_data(c, fr) for c in ("Ez", "Hy")
                      for fr in FREQS]) for b in built]


def test_batch_launch_shapes_and_autotune_are_bitwise(mi355x):
    """Every launch shape (rows per tile, waves per workgroup, workgroup
    order) gives the same bits, and so does the autotuner's pick; the tuner
    leaves the engine at the zero state a run starts from."""
    from prismo_amd import autotune
    from prismo_amd.core.batch import BatchEngine

    ref_eng, ref = _run(np.float32, "fp16", 130, True, chunk=8)
    ref_f = [ref_eng.fields(d) for d in range(len(SIZES))]
    ref_s = _spectra(ref)
    runs = []
    for chunk, waves, xcd in ((4, 8, 0), (16, 2, 1), (32, 1, 0)):
        built = [_design(nxp, nyp, 8, np.float32, 100 + k) for k, (nxp, nyp) in enumerate(SIZES)]
        eng = BatchEngine([b[0] for b in built], dtype=np.float32, coefficient_storage="fp16",
                          chunk=chunk, sample_T=32)
        eng.ctx.set_tuning(waves, xcd)
        eng.run_steps(130, graph=True)
        runs.append((eng, built))
    autotune._CACHE.clear()
    built = [_design(nxp, nyp, 8, np.float32, 100 + k) for k, (nxp, nyp) in enumerate(SIZES)]
    eng = BatchEngine([b[0] for b in built], dtype=np.float32, coefficient_storage="fp16",
                      chunk="auto", sample_T=32)
    eng.run_steps(130, graph=True)
    print("batch tuning:", eng.tuning)
    assert eng.tuning and not eng.tuning["cached"]
    assert eng.tuning["chunk"] in autotune.BATCH_CANDIDATES["chunk"]
    runs.append((eng, built))
    for eng, built in runs:
        s = _spectra(built)
        for d in range(len(SIZES)):
            f = eng.fields(d)
            for c in ("Ez", "Hx", "Hy"):
                assert np.array_equal(f[c], ref_f[d][c]), (d, c)
            assert np.array_equal(s[d], ref_s[d]), d
    # a second engine of the same batch size takes the cached pick
    built = [_design(nxp, nyp, 8, np.float32, 100 + k) for k, (nxp, nyp) in enumerate(SIZES)]
    eng2 = BatchEngine([b[0] for b in built], dtype=np.float32, coefficient_storage="fp16",
                       chunk="auto", sample_T=32)
    eng2.run_steps(10, graph=False)
    assert eng2.tuning["cached"]
