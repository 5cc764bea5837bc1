"""GPU: the dispersive-waveguide workload (C3) as assembled, against O1.

workloads.waveguide_3d on an 80 x 64 x 56-cell grid (12-cell CPML, the
450 x 220 nm Lorentz-silicon core on the SiO2 substrate painted on the
device, the Drude pad above it, the grid-mode TF/SF source, the 64-frequency
Ey/Ez/Hy/Hz plane) for 450 steps, against O1 built independently: node
permittivities from O3 (oracle/geometry.py) on the same shapes, Cb averaged
to the E points as solver.py:458-533 does, the two dispersive boxes added as
ADE regions with the coefficients of src/prismo/materials/dispersion.py:
189-231, 267-286, the same mode source's host update as the source, RegionDFT
as the monitor.  Bars: FP64 1e-10 relative L2 on the six fields and on the
spectra; FP32 fields with FP16 Cb 1e-5 against O1 fed the same FP16-rounded
coefficients (arithmetic error only).
"""

import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
DIMS = (80, 64, 56)
STEPS = 450
EPS0 = 8.854187817e-12


def _oracle(sim, fp16):
    from oracle import compat as O0
    from oracle import geometry as G
    from oracle.yee import YeeOracle

    g = sim.grid
    nx, ny, nz = g.dimensions
    ax = [g.origin[a] + np.arange(n) * g.spacing[a] for a, n in enumerate(g.dimensions)]
    shapes, lut = [], [1.0]
    for k, s in enumerate(sim.wg_shapes):
        shapes.append({"kind": "box", "center": s.center, "size": s.size, "material": k + 1})
        lut.append(s.material.epsilon_r)
    eps = np.asarray(lut)[G.paint(shapes, *ax)]
    shp = (nx, ny, nz)
    _, Cb, _, Db = O0.coefficients(eps, np.ones(shp), np.zeros(shp), np.zeros(shp), sim.dt)
    cb = {"Ex": O0._ex(Cb), "Ey": O0._ey(Cb), "Ez": O0._ez(Cb)}
    if fp16:
        scale = sim.dt / EPS0
        cb = {c: (v / scale).astype(np.float16).astype(np.float64) * scale for c, v in cb.items()}
    o1 = YeeOracle(nx, ny, nz, g.spacing, sim.dt, cb, float(Db.flat[0]), pml=(12,) * 6)
    for r in sim._ade:
        o1.add_ade(r.comp, r.box, r.kind, r.eps_inf, r.c)
    return o1


@pytest.mark.parametrize("dtype,storage,bar", [(np.float64, "auto", 1e-10),
                                               (np.float32, "fp16", 1e-5)])
def test_waveguide_reduced_vs_oracle(mi355x, dtype, storage, bar):
    from oracle.compat import RegionDFT
    from prismo_amd import workloads as W

    sim, mon, mode, _ = W.waveguide_3d(dtype=dtype, coefficient_storage=storage, dims=DIMS,
                                       pulse_width=3e-15)
    assert len(sim._ade) == 6 and {r.kind for r in sim._ade} == {1, 2}
    sim.run_steps(STEPS)
    src = sim.sources[0]
    o1 = _oracle(sim, storage == "fp16")
    ref = RegionDFT({c: mon._monitor_region[c] for c in mon.components}, mon.frequencies)
    o1.run(STEPS, sources=[lambda F_, t: src.update_fields(F_, t, sim.dt)], monitors=[ref])
    comps = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
    got = np.concatenate([sim.fields.to_numpy(c).ravel() for c in comps])
    want = np.concatenate([o1.F[c].ravel() for c in comps])
    err = rel_l2(got, want)
    spec_got = np.stack([np.stack([mon.get_frequency_data(c, f) for f in mon.frequencies])
                         for c in mon.components])
    spec_ref = np.stack([np.stack([ref.data[c][f] for f in mon.frequencies])
                         for c in mon.components])
    serr = rel_l2(spec_got, spec_ref)
    print(f"waveguide {np.dtype(dtype).name}/{storage}: fields {err:.3e}, spectra {serr:.3e}, "
          f"grid n_eff {src.grid_neff:.4f}")
    assert np.max(np.abs(want)) > 0 and np.max(np.abs(spec_ref)) > 0
    assert err <= bar
    assert serr <= bar
