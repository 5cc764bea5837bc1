"""GPU: the block DFT on the matrix cores (csrc/dft_mfma.hip).

The device spectra of a FieldMonitor are checked against the running sum of
src/prismo/monitors/field.py:138-143 evaluated in FP64 NumPy over the very
samples the device recorded (time_domain=True gives them bit for bit), so
only the spectral arithmetic is under test: the MFMA block products, the
frequency and column tails (counts not multiples of 16), runs that end
mid-block, and the compensated FP32 accumulation over many blocks.
Bars: FP64 1e-12 relative L2; FP32 1e-5 (the precision form's bound
against the FP64 result).
"""

import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
D = 20e-9


def _sim(dtype, nf, steps, comps=("Ez", "Hy", "Ex")):
    import prismo_amd as pa

    n, t = 36, 5
    rng = np.random.default_rng(11)
    eps = rng.choice([1.0, 2.085, 12.1], size=(n, n, 24))
    sim = pa.Simulation(size=((n - 2 * t) * D * (1 - 1e-9), (n - 2 * t) * D * (1 - 1e-9),
                              (24 - 2 * t) * D * (1 - 1e-9)),
                        resolution=1 / D, pml_layers=t, courant_factor=0.99, scheme="yee",
                        pml=pa.PMLParams(t), material_arrays={"eps_rel": eps}, dtype=dtype)
    sim.add_source(pa.ElectricDipole(position=(8 * D, 9 * D, 6 * D), polarization="z",
                                     frequency=193e12, pulse_width=4e-15))
    freqs = list(np.linspace(180e12, 210e12, nf))
    mon = pa.FieldMonitor(center=(14 * D, 12 * D, 7 * D), size=(0, 17 * D, 9 * D),
                          components=list(comps), frequencies=freqs, time_domain=True)
    sim.add_monitor(mon)
    sim.run_steps(steps)
    return sim, mon


def _fp64_dft(mon, c, f, dt):
    t, x = mon.get_time_data(c)
    ph = np.exp(-1j * 2 * np.pi * f * t) * dt
    return np.tensordot(ph, x.astype(np.float64), axes=(0, 0))


@pytest.mark.parametrize("nf", [64, 20, 11, 3])
@pytest.mark.parametrize("steps", [333, 128])
def test_dft_fp64_matches_running_sum(mi355x, nf, steps):
    sim, mon = _sim(np.float64, nf, steps)
    for c in mon.components:
        got = np.stack([mon.get_frequency_data(c, f) for f in mon.frequencies])
        ref = np.stack([_fp64_dft(mon, c, f, sim.dt) for f in mon.frequencies])
        assert rel_l2(got, ref) <= 1e-12, (c, nf, steps)


def test_dft_fp32_compensated_long_run(mi355x):
    """1500 steps = 12 blocks of 128 rows; 64 frequencies."""
    sim, mon = _sim(np.float32, 64, 1500, comps=("Ez", "Hy"))
    for c in mon.components:
        got = np.stack([mon.get_frequency_data(c, f) for f in mon.frequencies])
        ref = np.stack([_fp64_dft(mon, c, f, sim.dt) for f in mon.frequencies])
        err = rel_l2(got, ref)
        assert err <= 1e-5, (c, err)


def test_dft_split_runs_equal_one_run(mi355x):
    """Runs split at arbitrary step counts accumulate the same spectra."""
    import prismo_amd as pa

    _, whole = _sim(np.float64, 16, 300)
    sim, mon = _sim(np.float64, 16, 0)
    for k in (77, 1, 130, 92):
        sim.run_steps(k)
    for c in mon.components:
        for f in mon.frequencies:
            assert rel_l2(mon.get_frequency_data(c, f), whole.get_frequency_data(c, f)) <= 1e-13
