"""CPU: DFTMonitor and GaussianBeamSource host forms.

* DFTMonitor (src/prismo/monitors/dft.py:108-160): the region form sums the
  monitor region of each component (the FieldMonitor semantics of
  field.py:138-143), the compat placeholder sums the reference's fixed
  field[:10, :10] corner, both equal to O0's RegionDFT over the same field
  sequence; the 3-D placeholder fails as the reference's does (its corner
  does not broadcast against the (10, 10) accumulator).
* GaussianBeamSource (src/prismo/sources/gaussian.py:167-298): the added
  values form exp(-r^2 / w0^2) times the waveform on the source plane: 1 at
  a centre that sits on a node, symmetric about it, neighbour ratios
  exp(-(2m + 1) d^2 / w0^2); H gets E / 377; a region that reaches past the
  physical domain stops at the physical count minus one (the clamp of
  grid.py:345-352); in 3-D the flat profile does not broadcast against the
  component region and the add raises, as in the reference.
"""

import numpy as np
import pytest

import prismo_amd as pa
from oracle import compat as O0

D = 20e-9
SHRINK = 1 - 1e-9


def _grid(nx, ny, nz, pml):
    size = (nx * D * SHRINK, ny * D * SHRINK, nz * D * SHRINK if nz else 0.0)
    return pa.YeeGrid(pa.GridSpec(size=size, resolution=1 / D, boundary_layers=pml))


def _fields(g, rng):
    nx, ny, nz = g.dimensions
    F = O0.zeros(nx, ny, nz, g.is_2d)
    for c in F:
        F[c][...] = rng.standard_normal(F[c].shape)
    return F


@pytest.mark.parametrize("placeholder", [False, True])
def test_dft_monitor_host_update(placeholder):
    pa.set_backend("numpy")
    g = _grid(30, 20, 0, 4)
    comps = ["Ez", "Hx", "Hy"]
    freqs = [186e12, 193e12, 204e12]
    mon = pa.DFTMonitor(center=(12 * D, 9 * D, 0.0), size=(6 * D, 8 * D, 0.0),
                        frequencies=freqs, components=comps, compat_placeholder=placeholder)
    mon.initialize(g)
    if placeholder:
        regions = {c: np.ix_(range(10), range(10)) for c in comps}
    else:
        regions = {c: mon._monitor_region[c] for c in comps}
    ref = O0.RegionDFT(regions, freqs)
    rng = np.random.default_rng(5)
    dt = 4.67e-17
    t = 0.0
    for _ in range(25):
        t += dt
        F = _fields(g, rng)
        mon.update(F, t, dt)
        ref.update(F, t, dt)
    assert mon._time_steps == 25
    for c in comps:
        got = mon.get_frequency_data(c)
        for i, f in enumerate(freqs):
            np.testing.assert_allclose(got[f], ref.data[c][f], rtol=1e-13, atol=0)
            np.testing.assert_allclose(mon.get_frequency_data(c, i), ref.data[c][f], rtol=1e-13)
        ps = np.array([np.sum(np.abs(ref.data[c][f]) ** 2) for f in freqs])
        np.testing.assert_allclose(mon.get_power_spectrum(c), ps / ps.max(), rtol=1e-12)
    if placeholder:
        assert mon.get_frequency_data("Ez", 0).shape == (10, 10)
    else:
        assert mon.get_frequency_data("Ez", 0).shape == tuple(ix.size for ix in regions["Ez"])
    with pytest.raises(ValueError):
        mon.get_frequency_data("Ex")


def test_dft_monitor_placeholder_3d_fails():
    pa.set_backend("numpy")
    g = _grid(12, 12, 12, 2)
    mon = pa.DFTMonitor(center=(0.0, 0.0, 0.0), size=(4 * D, 4 * D, 4 * D),
                        frequencies=[193e12], compat_placeholder=True)
    with pytest.raises(ValueError):
        mon.initialize(g)


def test_gaussian_beam_profile_2d():
    pa.set_backend("numpy")
    g = _grid(40, 30, 0, 5)
    yc = 14 * D                       # node j = 14 + 5
    w0 = 4 * D
    beam = pa.GaussianBeamSource(center=(6 * D, yc, 0.0), size=(0.0, 12 * D, 0.0),
                                 direction="x", polarization="z", frequency=193e12,
                                 beam_waist=w0, pulse_width=4e-15)
    beam.initialize(g)
    nx, ny, nz = g.dimensions
    F = O0.zeros(nx, ny, nz, True)
    t = 23e-15
    amp = beam.waveform(t)
    beam.update_fields(F, t, 3e-17)
    jc = 19
    ez = F["Ez"]
    rows = np.nonzero(np.any(ez != 0, axis=1))[0]
    cols = np.nonzero(np.any(ez != 0, axis=0))[0]
    assert list(rows) == [11]                      # x = 6 d -> node 11, the source plane
    assert cols[0] == jc - 6 and cols[-1] == jc + 5  # [jc - 6, jc + 6): 12 nodes
    prof = ez[11, cols] / amp
    assert prof[6] == pytest.approx(1.0, abs=1e-15)              # the centre node
    np.testing.assert_allclose(prof[1:6], prof[7:12][::-1], rtol=1e-13)  # symmetric
    for m in range(5):                                           # neighbour ratios
        assert prof[6 + m + 1] / prof[6 + m] == pytest.approx(
            np.exp(-(2 * m + 1) * D**2 / w0**2), rel=1e-12)
    np.testing.assert_allclose(F["Hy"][11, cols], ez[11, cols] / 377.0, rtol=1e-15)
    assert not any(np.any(F[c]) for c in ("Ex", "Ey", "Hx", "Hz"))


def test_gaussian_beam_region_clamped_at_physical_edge():
    pa.set_backend("numpy")
    g = _grid(40, 30, 0, 5)
    beam = pa.GaussianBeamSource(center=(6 * D, 27 * D, 0.0), size=(0.0, 12 * D, 0.0),
                                 direction="x", polarization="z", frequency=193e12,
                                 beam_waist=4 * D, pulse_width=4e-15)
    beam.initialize(g)
    nx, ny, nz = g.dimensions
    F = O0.zeros(nx, ny, nz, True)
    beam.update_fields(F, 23e-15, 3e-17)
    cols = np.nonzero(F["Ez"][11])[0]
    # upper corner 33 d -> node 38, clamped to the physical count minus one
    # (30 - 1 = 29): the region is [26, 29)
    assert cols[0] == 27 + 5 - 6 and cols[-1] == 28


def test_gaussian_beam_3d_add_does_not_broadcast():
    pa.set_backend("numpy")
    g = _grid(10, 10, 10, 2)
    beam = pa.GaussianBeamSource(center=(0.0, 0.0, 4 * D), size=(4 * D, 4 * D, 0.0),
                                 direction="z", polarization="x", frequency=193e12,
                                 beam_waist=2 * D, pulse_width=4e-15)
    beam.initialize(g)
    nx, ny, nz = g.dimensions
    F = O0.zeros(nx, ny, nz, False)
    with pytest.raises(ValueError):
        beam.update_fields(F, 23e-15, 3e-17)
