"""GPU parity: libpmfdtd through the prismo_amd API against the oracles.

compat scheme vs O0 (oracle/compat.py): bit-identical fields and time
samples, spectra within 1e-10 relative L2 (the device sums each block of
steps as one product instead of one step at a time).
yee scheme vs O1 (oracle/yee.py): FP64 within 1e-10 relative L2 on fields
and spectra; FP32 within the bounds stated per test.
"""

import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

SEED = 20261202
D = 20e-9


def _grid2d(nxp, nyp, pml, cf=0.99):
    import prismo_amd as pa

    return dict(size=(nxp * D * (1 - 1e-9), nyp * D * (1 - 1e-9), 0.0), resolution=1 / D,
                pml_layers=pml, courant_factor=cf)


def _same_bits(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def _o0_from_sim(sim, eps, sig):
    from oracle import compat as O0

    nx, ny, nz = sim.grid.dimensions
    shp = (nx, ny, nz)
    C = dict(zip(("Ca", "Cb", "Da", "Db"),
                 O0.coefficients(eps.reshape(shp), np.ones(shp), sig.reshape(shp),
                                 np.zeros(shp), sim.dt)))
    F = O0.zeros(nx, ny, nz, sim.grid.is_2d)
    return F, C


@pytest.mark.parametrize("lossy", [False, True])
def test_compat_2d_bitwise(mi355x, lossy):
    import prismo_amd as pa
    from oracle import compat as O0

    rng = np.random.default_rng(SEED)
    kw = _grid2d(40, 28, 6)
    g = pa.YeeGrid(pa.GridSpec(kw["size"], kw["resolution"], kw["pml_layers"]))
    nx, ny, _ = g.dimensions
    eps = rng.uniform(1.0, 12.1, size=(nx, ny, 1))
    sig = (np.where(rng.uniform(size=(nx, ny, 1)) < 0.3, rng.uniform(0, 3e4, size=(nx, ny, 1)), 0.0)
           if lossy else np.zeros((nx, ny, 1)))
    sim = pa.Simulation(**kw, material_arrays={"eps_rel": eps, "sigma_e": sig})
    src = pa.PlaneWaveSource(center=(5 * D, 12 * D, 0), size=(0, 14 * D, 0), direction="+x",
                             polarization="z", frequency=193e12, pulse_width=10e-15)
    dip = pa.ElectricDipole(position=(30 * D, 6 * D, 0), polarization="z", frequency=200e12,
                            pulse_width=5e-15, amplitude=0.5, phase=0.3)
    mon = pa.FieldMonitor(center=(33 * D, 12 * D, 0), size=(0, 14 * D, 0),
                          components=["Ez", "Hy", "Hx"],
                          frequencies=[187e12, 193e12, 200e12], time_domain=True)
    sources = [src, dip]
    if lossy:
        # a second dipole on the same Ez cell: overlapping boxes of one
        # component are added box by box, in order (the others in one launch)
        sources.append(pa.ElectricDipole(position=(30 * D, 6 * D, 0), polarization="z",
                                         frequency=187e12, pulse_width=7e-15, amplitude=0.3))
    for s in sources:
        sim.add_source(s)
    sim.add_monitor(mon)
    sim.run(150 * sim.dt * 0.9999)  # ceil -> 150 steps
    assert sim.step_count == 150

    F, C = _o0_from_sim(sim, eps, sig)
    ref_mon = O0.RegionDFT({c: mon._monitor_region[c] for c in mon.components}, mon.frequencies)
    srcs = [lambda F_, t, s=s: s.update_fields(F_, t, sim.dt) for s in sources]
    t = O0.run(F, C, g.spacing, True, sim.dt, 150, sources=srcs, monitors=[ref_mon])
    assert t == sim.current_time
    for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        got = sim.fields.to_numpy(c)
        assert _same_bits(got, F[c]), f"{c}: max diff {np.nanmax(np.abs(got - F[c]))}"
    for c in mon.components:
        td = mon.get_time_data(c)[1]
        assert _same_bits(td, np.array(ref_mon.samples[c])), c
        for f in mon.frequencies:
            assert rel_l2(mon.get_frequency_data(c, f), ref_mon.data[c][f]) <= 1e-10


def test_compat_2d_te_gates(mi355x):
    """Seeded TE fields exercise both 2-D gates and the TE E update."""
    import prismo_amd as pa
    from oracle import compat as O0

    rng = np.random.default_rng(SEED + 1)
    kw = _grid2d(30, 22, 3, cf=0.9)
    sim = pa.Simulation(**kw)
    nx, ny, _ = sim.grid.dimensions
    ex0 = rng.standard_normal(sim.fields["Ex"].shape)
    ey0 = rng.standard_normal(sim.fields["Ey"].shape)
    sim.fields["Ex"] = ex0
    sim.fields["Ey"] = ey0
    sim.add_source(pa.MagneticDipole(position=(15 * D, 11 * D, 0), polarization="z",
                                     frequency=150e12, pulse=False, amplitude=0.01))
    sim.run(40 * sim.dt * 0.9999)
    F, C = _o0_from_sim(sim, np.ones((nx, ny, 1)), np.zeros((nx, ny, 1)))
    F["Ex"][:] = ex0
    F["Ey"][:] = ey0
    O0.run(F, C, sim.grid.spacing, True, sim.dt, 40,
           sources=[lambda F_, t: sim.sources[0].update_fields(F_, t, sim.dt)])
    for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        assert _same_bits(sim.fields.to_numpy(c), F[c]), c


def test_compat_3d_bitwise(mi355x):
    import prismo_amd as pa
    from oracle import compat as O0

    rng = np.random.default_rng(SEED + 2)
    n = 14
    sim = pa.Simulation(size=((n - 4) * D * (1 - 1e-9),) * 3, resolution=1 / D, pml_layers=2,
                        courant_factor=0.99,
                        material_arrays={"eps_rel": rng.choice([1.0, 2.085, 12.1], size=(n, n, n))})
    eps = sim.solver.updater.eps_rel
    sim.add_source(pa.ElectricDipole(position=(5 * D, 5 * D, 5 * D), polarization="z",
                                     frequency=193e12, pulse_width=10e-15))
    mon = pa.FieldMonitor(center=(7 * D, 5 * D, 5 * D), size=(0, 4 * D, 4 * D),
                          components=["Ez", "Ex"], frequencies=[190e12, 195e12],
                          time_domain=False)
    sim.add_monitor(mon)
    sim.run(30 * sim.dt * 0.9999)
    F, C = _o0_from_sim(sim, eps, np.zeros_like(eps))
    ref_mon = O0.RegionDFT({c: mon._monitor_region[c] for c in mon.components}, mon.frequencies)
    O0.run(F, C, sim.grid.spacing, False, sim.dt, 30,
           sources=[lambda F_, t: sim.sources[0].update_fields(F_, t, sim.dt)],
           monitors=[ref_mon])
    for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        assert _same_bits(sim.fields.to_numpy(c), F[c]), c
    for c in mon.components:
        for f in mon.frequencies:
            assert rel_l2(mon.get_frequency_data(c, f), ref_mon.data[c][f]) <= 1e-10


def test_updater_half_steps_match_oracle(mi355x):
    """MaxwellUpdater.update_magnetic_fields / update_electric_fields alone."""
    import prismo_amd as pa
    from oracle import compat as O0

    rng = np.random.default_rng(SEED + 3)
    g = pa.YeeGrid(pa.GridSpec((9 * D * (1 - 1e-9), 7 * D * (1 - 1e-9), 5 * D * (1 - 1e-9)),
                               1 / D, 1))
    dt = g.suggest_time_step(0.5)
    up = pa.MaxwellUpdater(g, dt)
    fl = pa.ElectromagneticFields(g)
    nx, ny, nz = g.dimensions
    F = O0.zeros(nx, ny, nz, False)
    for c in F:
        F[c][:] = rng.standard_normal(F[c].shape)
        fl[c] = F[c]
    C = dict(Ca=up.Ca, Cb=up.Cb, Da=up.Da, Db=up.Db)
    up.update_magnetic_fields(fl)
    O0.h3d(F, C, g.spacing)
    for c in F:
        assert _same_bits(fl.to_numpy(c), F[c]), c
    up.update_electric_fields(fl)
    O0.e3d(F, C, g.spacing)
    for c in F:
        assert _same_bits(fl.to_numpy(c), F[c]), c


# ----------------------------------------------------------------- yee scheme
def _yee_oracle(sim, eps, pml_t, dtype=np.float64, cb_override=None):
    from oracle import compat as O0
    from oracle.yee import YeeOracle

    g = sim.grid
    nx, ny, nz = g.dimensions
    shp = (nx, ny, nz)
    Ca, Cb, Da, Db = O0.coefficients(eps.reshape(shp), np.ones(shp), np.zeros(shp),
                                     np.zeros(shp), sim.dt)
    if g.is_2d:
        cb = {"Ez": O0._ez2(Cb[:, :, 0])}
    else:
        cb = {"Ex": O0._ex(Cb), "Ey": O0._ey(Cb), "Ez": O0._ez(Cb)}
    if cb_override is not None:
        cb = cb_override
    pml = None
    if pml_t:
        pml = (pml_t,) * 4 + ((0, 0) if g.is_2d else (pml_t, pml_t))
    return YeeOracle(nx, ny, nz, g.spacing, sim.dt, cb, float(Db.flat[0]), pml=pml, dtype=dtype)


def _run_o1(o1, sim, nsteps, mon=None):
    from oracle.compat import RegionDFT

    ref = RegionDFT({c: mon._monitor_region[c] for c in mon.components}, mon.frequencies) if mon else None
    t = 0.0
    for _ in range(nsteps):
        o1.step()
        t += sim.dt
        for s in sim.sources:
            s.update_fields(o1.F, t, sim.dt)
        if ref:
            ref.update(o1.F, t, sim.dt)
    return ref


def test_yee_2d_cpml_fp64(mi355x):
    import prismo_amd as pa

    rng = np.random.default_rng(SEED + 4)
    kw = _grid2d(48, 30, 8)
    g = pa.YeeGrid(pa.GridSpec(kw["size"], kw["resolution"], kw["pml_layers"]))
    nx, ny, _ = g.dimensions
    eps = np.ones((nx, ny, 1))
    eps[:, ny // 2 - 3:ny // 2 + 3] = 12.1
    eps *= rng.uniform(1.0, 1.2, size=eps.shape)
    sim = pa.Simulation(**kw, scheme="yee", pml=pa.PMLParams(8),
                        material_arrays={"eps_rel": eps})
    sim.add_source(pa.PlaneWaveSource(center=(12 * D, 15 * D, 0), size=(0, 10 * D, 0),
                                      direction="+x", polarization="z", frequency=193e12,
                                      pulse_width=5e-15))
    mon = pa.FieldMonitor(center=(40 * D, 15 * D, 0), size=(0, 10 * D, 0),
                          components=["Ez", "Hy"], frequencies=[187e12, 193e12, 200e12],
                          time_domain=False)
    sim.add_monitor(mon)
    n = 400
    sim.run(n * sim.dt * 0.9999)
    o1 = _yee_oracle(sim, eps, 8)
    ref = _run_o1(o1, sim, n, mon)
    for c in ("Ez", "Hx", "Hy"):
        assert rel_l2(sim.fields.to_numpy(c), o1.F[c]) <= 1e-10, c
    for c in mon.components:
        for f in mon.frequencies:
            assert rel_l2(mon.get_frequency_data(c, f), ref.data[c][f]) <= 1e-10


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_yee_3d_cpml(mi355x, dtype):
    import prismo_amd as pa

    rng = np.random.default_rng(SEED + 5)
    n, t = 26, 5
    eps_blocks = rng.choice([1.0, 2.085, 12.1], size=(4, 4, 4))
    eps = np.kron(eps_blocks, np.ones((8, 8, 8)))[:n, :n, :n]
    sim = pa.Simulation(size=((n - 2 * t) * D * (1 - 1e-9),) * 3, resolution=1 / D,
                        pml_layers=t, courant_factor=0.99, scheme="yee", pml=pa.PMLParams(t),
                        material_arrays={"eps_rel": eps}, dtype=getattr(np, dtype))
    sim.add_source(pa.ElectricDipole(position=(8 * D, 8 * D, 8 * D), polarization="z",
                                     frequency=193e12, pulse_width=5e-15))
    mon = pa.FieldMonitor(center=(12 * D, 8 * D, 8 * D), size=(0, 6 * D, 6 * D),
                          components=["Ez", "Ex", "Hy"], frequencies=[190e12, 196e12],
                          time_domain=False)
    sim.add_monitor(mon)
    steps = 200
    sim.run(steps * sim.dt * 0.9999)
    o1 = _yee_oracle(sim, eps, t)
    ref = _run_o1(o1, sim, steps, mon)
    # FP32 fields against the FP64 oracle fed the same FP64 coefficients:
    # the FP32 accuracy bar, 1e-5, on every component and every spectrum
    tol = 1e-10 if dtype == "float64" else 1e-5
    comps = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
    errs = {c: rel_l2(sim.fields.to_numpy(c), o1.F[c]) for c in comps}
    whole = rel_l2(np.concatenate([sim.fields.to_numpy(c).ravel() for c in comps]),
                   np.concatenate([o1.F[c].ravel() for c in comps]))
    spec = {(c, f): rel_l2(mon.get_frequency_data(c, f), ref.data[c][f])
            for c in mon.components for f in mon.frequencies}
    print(f"yee 3-D {dtype}: six-component {whole:.3e}; per component "
          + " ".join(f"{c} {e:.2e}" for c, e in errs.items())
          + "; spectra max " + f"{max(spec.values()):.2e}")
    assert whole <= tol
    for c, e in errs.items():
        assert e <= tol, (c, e)
    for k, e in spec.items():
        assert e <= tol, (k, e)
