"""GPU: per-design failure in a batched sweep.

One design's permittivity is poisoned with a NaN on the host (its Cb at
that cell is then NaN, so its Ez leaves the finite numbers on the first
step).  The kernel flags that design alone (``pm_batch_design_status``); the
sweep reports it with status 'error' and the other designs' spectra are
bit-identical to those of the same batch without the poison (no design's
arithmetic reads another's memory, and the spectral GEMM keeps columns
apart).  A design the batch cannot take (a time-domain monitor) runs alone
through its own Simulation, and its results equal that Simulation run by
itself.
"""

import numpy as np
import pytest

from tests.test_gpu_batch import FREQS, SIZES, _design

pytestmark = pytest.mark.gpu
STEPS = 120


def _build(p):
    sim, _ = _design(*SIZES[p["k"]], 8, p["dtype"], 100 + p["k"], poison=p.get("poison"))
    if p.get("time_domain"):
        import prismo_amd as pa

        D = 20e-9
        sim.add_monitor(pa.FieldMonitor(center=(20 * D, 10 * D, 0), size=(0, 4 * D, 0),
                                        components=["Ez"], frequencies=FREQS,
                                        time_domain=True))
    return sim


def _analyze(sim, p):
    mon = sim.monitors[0]
    return {"spec": np.stack([mon.get_frequency_data(c, f) for c in mon.components
                              for f in FREQS])}


def _sweep(samples, dtype, storage):
    from prismo_amd.optimization.sweep import BatchedRunner, ParameterSweep

    sw = ParameterSweep.from_samples(
        samples, batched=BatchedRunner(build=_build, analyze=_analyze, steps=STEPS,
                                       dtype=dtype, coefficient_storage=storage,
                                       engine_kw={"sample_T": 32, "chunk": 32}))
    return sw.run()


@pytest.mark.parametrize("dtype,storage", [(np.float32, "fp16"), (np.float64, "auto")])
def test_poisoned_design_fails_alone(mi355x, dtype, storage):
    clean = [{"k": k, "dtype": dtype} for k in range(len(SIZES))]
    bad = [dict(c, poison=(c["k"] == 2)) for c in clean]
    ref = _sweep(clean, dtype, storage)
    got = _sweep(bad, dtype, storage)
    print("status:", [r["status"] for r in got])
    assert [r["status"] for r in ref] == ["success"] * len(SIZES)
    assert [r["status"] for r in got] == ["success", "success", "error", "success"]
    assert "diverged" in got[2]["error"]
    for k in (0, 1, 3):
        assert np.array_equal(got[k]["spec"], ref[k]["spec"]), k


def test_engine_reports_diverged_designs(mi355x):
    from prismo_amd.core.batch import BatchEngine

    sims = [_build({"k": k, "dtype": np.float32, "poison": k == 1}) for k in range(len(SIZES))]
    eng = BatchEngine(sims, dtype=np.float32, coefficient_storage="auto", chunk=32,
                      sample_T=32)
    eng.run_steps(40)
    assert eng.diverged == [False, True, False, False]
    assert np.isfinite(eng.fields(0)["Ez"]).all()
    assert not np.isfinite(eng.fields(1)["Ez"]).all()


def test_unbatchable_design_runs_alone(mi355x):
    samples = [{"k": k, "dtype": np.float32} for k in range(3)]
    samples[1]["time_domain"] = True
    got = _sweep(samples, np.float32, "auto")
    assert [r["status"] for r in got] == ["success"] * 3
    alone = _build(samples[1])
    alone.run_steps(STEPS)
    assert np.array_equal(got[1]["spec"], _analyze(alone, samples[1])["spec"])
    # the batched designs match a batch of just those two
    two = _sweep([samples[0], samples[2]], np.float32, "auto")
    assert np.array_equal(got[0]["spec"], two[0]["spec"])
    assert np.array_equal(got[2]["spec"], two[1]["spec"])
