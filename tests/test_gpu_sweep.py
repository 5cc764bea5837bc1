"""GPU: coupler sweeps through ParameterSweep's batched back end against
the plain FP64 path (each design's own Simulation.run on the step library).

* batched FP64 equals sequential FP64 (two different kernels: the batched
  ping-pong tile kernel and the per-grid 2-D stencil) to 1e-9 in |S21|,
  |S31| and port power, run as three pipelined groups (2 + 2 + 1 designs,
  each group queued while the previous one runs);
* batched FP32 fields with FP16 coefficients stay within 0.01 of the FP64
  |S21| and |S31| (the sweep's accuracy bar), on designs of different size.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
STEPS = 1200


def _designs():
    from prismo_amd import workloads as W

    ds = W.coupler_designs(64)
    # small and uneven: shortest / narrowest few of the seeded set
    ds = sorted(ds, key=lambda d: W.coupler_cells(d)[0] * W.coupler_cells(d)[1])[:5]
    return ds


def _sequential_fp64(ds):
    from prismo_amd import workloads as W

    out = []
    for p in ds:
        sim = W.coupler_sim(p, dtype=np.float64)
        sim.run_steps(STEPS)
        out.append(W.coupler_analyze(sim))
    return out


def _batched(ds, dtype, storage, max_batch=256):
    from prismo_amd import workloads as W
    from prismo_amd.optimization.sweep import BatchedRunner, ParameterSweep

    sw = ParameterSweep.from_samples(
        ds, batched=BatchedRunner(build=lambda p: W.coupler_sim(p, backend=W.host_backend()),
                                  analyze=W.coupler_analyze, steps=STEPS, dtype=dtype,
                                  coefficient_storage=storage, max_batch=max_batch))
    res = sw.run()
    assert all(r["status"] == "success" for r in res), [r.get("error") for r in res]
    return res


def test_sweep_batched_fp64_equals_sequential(mi355x):
    ds = _designs()
    seq = _sequential_fp64(ds)
    bat = _batched(ds, np.float64, "auto", max_batch=2)
    for a, b in zip(seq, bat):
        for k in ("s21", "s31", "p_in"):
            np.testing.assert_allclose(b[k], a[k], rtol=1e-9, atol=1e-12 * np.max(np.abs(a[k])))


def test_sweep_fp32_fp16_within_one_percent(mi355x):
    ds = _designs()
    seq = _sequential_fp64(ds)
    bat = _batched(ds, np.float32, "fp16")
    for a, b in zip(seq, bat):
        for k in ("s21", "s31"):
            assert np.max(np.abs(b[k] - a[k])) <= 0.01, (k, np.max(np.abs(b[k] - a[k])))
