"""Mode source (total-field / scattered-field sheet of a solved mode) in a
silicon strip on oxide, painted on the device: the mode leaves the source
plane in the chosen direction only and keeps its shape downstream."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

D = 20e-9
F0 = 193e12


def _run(direction):
    import prismo_amd as pa

    si, ox = pa.Material("Si", 11.7), pa.Material("SiO2", 2.085)
    size = (3.2e-6 * (1 - 1e-9), 1.6e-6 * (1 - 1e-9), 1.2e-6 * (1 - 1e-9))
    sim = pa.Simulation(size=size, resolution=1 / D, pml_layers=10, courant_factor=0.99,
                        scheme="yee", dtype=np.float32, pml=pa.PMLParams(10))
    # physical domain [0, L] per axis (the CPML sits below 0 and above L)
    shapes = [pa.Box(ox, (1.6e-6, 0.8e-6, 0.0), (8e-6, 8e-6, 1.0e-6)),
              pa.Box(si, (1.6e-6, 0.8e-6, 0.61e-6), (8e-6, 0.46e-6, 0.22e-6))]
    for s in shapes:
        sim.add_structure(s)
    g = sim.grid
    nx, ny, nz = g.dimensions
    y = g.origin[1] + np.arange(ny) * g.dy
    z = g.origin[2] + np.arange(nz) * g.dz
    Y, Z = np.meshgrid(y, z, indexing="ij")
    pts = np.column_stack([np.zeros(Y.size), Y.ravel(), Z.ravel()])
    eps = np.ones(Y.size)
    for s in shapes:
        eps[s.contains(pts)] = s.material.epsilon_r
    mode = pa.ModeSolver(299792458.0 / F0, y, z, eps.reshape(Y.shape)).solve(1, "TE")[0]
    assert 2.0 < mode.neff.real < 3.0
    wf = pa.sources.waveform.GaussianPulse(F0, 10e-15)
    sim.add_source(pa.ModeSource(center=(0.8e-6, 0.8e-6, 0.6e-6), size=(0.0, 1.4e-6, 1.0e-6),
                                 mode=mode, direction=direction, waveform=wf))
    mons = {}
    for key, x in (("up", 0.4e-6), ("down", 2.6e-6)):
        mons[key] = pa.FieldMonitor(center=(x, 0.8e-6, 0.6e-6), size=(0.0, 1.4e-6, 1.0e-6),
                                    components=["Ey", "Ez", "Hy", "Hz"], frequencies=[F0],
                                    time_domain=False)
        sim.add_monitor(mons[key])
    sim.run(3000 * sim.dt * 0.9999)
    out = {}
    for key, m in mons.items():
        f = {c: np.squeeze(m.get_frequency_data(c, F0)) for c in ("Ey", "Ez", "Hy", "Hz")}
        s0 = min(a.shape[0] for a in f.values())
        s1 = min(a.shape[1] for a in f.values())
        f = {c: a[:s0, :s1] for c, a in f.items()}
        out[key] = (0.5 * np.real(f["Ey"] * np.conj(f["Hz"]) - f["Ez"] * np.conj(f["Hy"])).sum(),
                    f["Ey"])
    return out, mode


def test_mode_source_is_one_way_and_keeps_its_shape(mi355x):
    fwd, mode = _run("+x")
    p_up, p_down = fwd["up"][0], fwd["down"][0]
    assert p_down > 0
    assert abs(p_up) < 0.03 * p_down, (p_up, p_down)
    # downstream Ey keeps the injected mode's profile (mode x = simulation y)
    ey = fwd["down"][1]
    ref = np.real(mode.Ex)
    # crop the mode to the monitor window: 1.4 x 1.0 um about (0.8, 0.6) um
    sel_y = np.abs(mode.x - 0.8e-6) <= 0.7e-6 + 1e-12
    sel_z = np.abs(mode.y - 0.6e-6) <= 0.5e-6 + 1e-12
    ref = ref[np.ix_(sel_y, sel_z)][: ey.shape[0], : ey.shape[1]]
    e = ey[: ref.shape[0], : ref.shape[1]]
    ov = abs(np.vdot(ref, e)) / (np.linalg.norm(ref) * np.linalg.norm(e))
    assert ov > 0.95, ov

    bwd, _ = _run("-x")
    assert bwd["up"][0] < 0
    assert abs(bwd["down"][0]) < 0.03 * abs(bwd["up"][0])
    assert abs(abs(bwd["up"][0]) / p_down - 1) < 0.1
