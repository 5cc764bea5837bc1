"""GPU: a source and a monitor with no device form (user subclasses, as
src/prismo/sources/base.py:141-157 and src/prismo/monitors/base.py:136-150
define them) run through the per-step path on the device field views and
give what the device path gives for the same source and probe."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
D = 20e-9
STEPS = 80


def _run(custom):
    import prismo_amd as pa
    from prismo_amd.monitors.base import Monitor

    s = 1 - 1e-9
    sim = pa.Simulation(size=(40 * D * s, 30 * D * s, 0.0), resolution=1 / D, pml_layers=6,
                        courant_factor=0.99, scheme="yee", pml=pa.PMLParams(6))
    src_cls = pa.ElectricDipole
    if custom:
        class HostDipole(pa.ElectricDipole):
            """The same dipole without a device form: update_fields only."""

            def device_terms(self, times):
                return None

        src_cls = HostDipole
    sim.add_source(src_cls(position=(12 * D, 15 * D, 0), polarization="z", frequency=193e12,
                           pulse_width=5e-15))
    if custom:
        class Probe(Monitor):
            """Ez at one point, read from the field view every step."""

            def __init__(self, center):
                super().__init__(center, (0.0, 0.0, 0.0))
                self.values = []

            def update(self, fields, time, dt):
                v = fields["Ez"][self._monitor_region["Ez"]]
                self.values.append(float(v.detach().cpu().numpy().ravel()[0]))

        probe = Probe((30 * D, 15 * D, 0.0))
    else:
        probe = pa.FieldMonitor(center=(30 * D, 15 * D, 0.0), size=(0.0, 0.0, 0.0),
                                components=["Ez"], time_domain=True)
    sim.add_monitor(probe)
    sim.run(STEPS * sim.dt * 0.9999)
    assert sim.step_count == STEPS
    if custom:
        trace = np.array(probe.values)
    else:
        trace = np.array([np.ravel(a)[0] for a in probe.get_time_data("Ez")[1]])
    return sim.fields.to_numpy("Ez"), trace


def test_host_source_and_monitor_match_device_path(mi355x):
    ez_dev, tr_dev = _run(False)
    ez_host, tr_host = _run(True)
    assert len(tr_host) == len(tr_dev) == STEPS
    scale = np.max(np.abs(ez_dev))
    assert scale > 0
    # the device tables and the per-step evaluation of the pulse may differ
    # in the last bit; nothing else differs
    assert np.max(np.abs(ez_host - ez_dev)) <= 1e-12 * scale
    assert np.max(np.abs(tr_host - tr_dev)) <= 1e-12 * max(np.max(np.abs(tr_dev)), 1e-300)
    assert np.max(np.abs(tr_dev)) > 0


def test_host_source_seen_by_device_monitor(mi355x):
    """A source with no device form and a device monitor at its cell: the
    monitor samples after the source (src/prismo/core/simulation.py:158-164),
    so its samples and spectrum equal O0's run with the same source, bit for
    bit (samples) and to 1e-12 (spectrum)."""
    import prismo_amd as pa
    from oracle import compat as O0

    s = 1 - 1e-9
    sim = pa.Simulation(size=(24 * D * s, 18 * D * s, 0.0), resolution=1 / D, pml_layers=4,
                        courant_factor=0.99)

    class HostDipole(pa.ElectricDipole):
        def device_terms(self, times):
            return None

    src = HostDipole(position=(10 * D, 8 * D, 0), polarization="z", frequency=193e12,
                     pulse_width=3e-15)
    sim.add_source(src)
    mon = pa.FieldMonitor(center=(10 * D, 8 * D, 0.0), size=(0.0, 0.0, 0.0), components=["Ez"],
                          time_domain=True, frequencies=[188e12, 199e12])
    sim.add_monitor(mon)
    steps = 60
    sim.run_steps(steps)
    assert src._source_box["Ez"] == mon._monitor_box["Ez"]
    nx, ny, nz = sim.grid.dimensions
    shp = (nx, ny, nz)
    C = dict(zip(("Ca", "Cb", "Da", "Db"),
                 O0.coefficients(np.ones(shp), np.ones(shp), np.zeros(shp), np.zeros(shp), sim.dt)))
    F = O0.zeros(nx, ny, nz, True)
    ref = O0.RegionDFT({"Ez": mon._monitor_region["Ez"]}, mon.frequencies)
    O0.run(F, C, sim.grid.spacing, True, sim.dt, steps,
           sources=[lambda F_, t: src.update_fields(F_, t, sim.dt)], monitors=[ref])
    got = np.array([np.ravel(a)[0] for a in mon.get_time_data("Ez")[1]])
    want = np.array([np.ravel(a)[0] for a in ref.samples["Ez"]])
    assert np.array_equal(got, want)
    assert np.max(np.abs(want)) > 0
    for f in mon.frequencies:
        g = np.ravel(mon.get_frequency_data("Ez", f))[0]
        w = np.ravel(ref.data["Ez"][f])[0]
        assert abs(g - w) <= 1e-12 * abs(w)
