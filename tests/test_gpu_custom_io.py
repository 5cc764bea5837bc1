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
