"""CPU: the package's host-side restatements against vectors written by the
reference's own modules (tests/golden/reference_vectors.npz, made by
tests/golden/make_golden.py from src/prismo/core/grid.py,
src/prismo/sources/waveform.py and src/prismo/materials/dispersion.py).

Grid rules, index maps and the ADE coefficient formulas are integer or
correctly rounded FP64 arithmetic and must match exactly; waveform values go
through libm exp/sin, compared at 1e-15 relative.
"""

from pathlib import Path

import numpy as np
import pytest

GOLD = np.load(Path(__file__).parent / "golden" / "reference_vectors.npz")
SPECS = {
    "C1": ((10e-6, 5e-6, 0.0), 50e6, 20, 0.99),
    "C1exact": ((500 * 20e-9 * (1 - 1e-9), 250 * 20e-9 * (1 - 1e-9), 0.0), 50e6, 20, 0.99),
    "C2": ((232 * 20e-9 * (1 - 1e-9),) * 3, 50e6, 12, 0.99),
    "T3": ((10.0, 5.0, 3.0), 20.0, 5, 0.9),
    "T2": ((8.0, 6.0, 0.0), 25.0, 8, 0.5),
}
COMPS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def _grid(tag):
    from prismo_amd.core.grid import GridSpec, YeeGrid

    size, res, pml, _ = SPECS[tag]
    return YeeGrid(GridSpec(size=size, resolution=res, boundary_layers=pml))


def _region(ix):
    out = []
    for a in range(3):
        if a >= len(ix):
            out += [0, 1]
            continue
        v = np.asarray(ix[a]).ravel()
        out += [int(v[0]) if v.size else -1, int(v.size)]
    return np.array(out)


@pytest.mark.parametrize("tag", list(SPECS))
def test_grid_rules(tag):
    g = _grid(tag)
    cf = SPECS[tag][3]
    assert tuple(g.dimensions) == tuple(GOLD[f"{tag}_dims"])
    assert tuple(g.physical_dimensions) == tuple(GOLD[f"{tag}_phys"])
    assert g.get_time_step(cf) == float(GOLD[f"{tag}_dt"])
    assert g.get_courant_number(g.get_time_step(cf)) == float(GOLD[f"{tag}_courant"])
    assert np.array_equal(np.asarray(g.origin, dtype=float), GOLD[f"{tag}_origin"])
    for c in COMPS:
        assert tuple(g.get_field_shape(c)) == tuple(GOLD[f"{tag}_shape_{c}"]), c


@pytest.mark.parametrize("tag", list(SPECS))
def test_point_to_index_and_regions(tag):
    g = _grid(tag)
    got = np.array([g.point_to_index(tuple(p)) for p in GOLD[f"{tag}_pts"]])
    assert np.array_equal(got, GOLD[f"{tag}_pidx"])
    for c in COMPS:
        regs = np.array([_region(g.get_component_indices(c, *map(int, b)))
                         for b in GOLD[f"{tag}_boxes"]])
        assert np.array_equal(regs, GOLD[f"{tag}_reg_{c}"]), c


def test_c1_line_source_and_monitor_regions():
    """The 2 um line source at x = 1 um and the monitor line at x = 8 um of
    the 500x250-cell strip workload land on the reference's cells."""
    g = _grid("C1exact")
    for tag, cx in (("src", 1e-6), ("mon", 8e-6)):
        x0, y0, z0 = g.point_to_index((cx, 2.5e-6 - 1e-6, 0.0))
        x1, y1, z1 = g.point_to_index((cx, 2.5e-6 + 1e-6, 0.0))
        if x0 == x1:
            x1 = x0 + 1
        for c in ("Ez", "Hy"):
            reg = _region(g.get_component_indices(c, x0, x1, y0, y1, z0, z1))
            assert np.array_equal(reg, GOLD[f"C1_{tag}_{c}"]), (tag, c)


def test_waveforms():
    from prismo_amd.sources import waveform as W

    ts = GOLD["W_t"]
    cases = {
        "gauss": W.GaussianPulse(193e12, 10e-15),
        "gauss_ph": W.GaussianPulse(150e12, 7e-15, 2.0, 0.4, 20e-15),
        "cw": W.ContinuousWave(193e12, 1.5, 0.2),
        "ricker": W.RickerWavelet(100e12, 1.2),
    }
    for k, w in cases.items():
        ref = GOLD[f"W_{k}"]
        got = w.table(ts)
        np.testing.assert_allclose(got, ref, rtol=1e-15, atol=1e-300, err_msg=k)


def test_dispersion_closed_forms_and_ade_coefficients():
    from prismo_amd.materials import dispersion as M

    w = GOLD["D_w"]
    si = M.LorentzMaterial(7.98, [M.LorentzPole(omega_0=1.1e16, delta_epsilon=3.69, gamma=1e13)])
    two = M.LorentzMaterial(2.0, [M.LorentzPole(2e15, 1.5, 3e13), M.LorentzPole(9e15, 0.7, 1e14)])
    dr = M.DrudeMaterial(1.0, 1.37e16, 1.07e14)
    db = M.DebyeMaterial(2.0, 5.0, 1e-14)
    for name, m in (("si", si), ("two", two), ("drude", dr), ("debye", db)):
        np.testing.assert_allclose(m.permittivity(w), GOLD[f"D_{name}_eps"], rtol=1e-15,
                                   err_msg=name)
    for dt in (3.8131497e-17, 4.67e-17, 1e-16):
        key = f"{dt:.3e}"
        for name, m in (("si", si), ("two", two)):
            co = m.get_ade_coefficients(dt)
            got = np.array([[p["C0"], p["C1"], p["C2"], p["C3"]] for p in co["poles"]])
            assert np.array_equal(got, GOLD[f"D_{name}_ade_{key}"]), (name, key)
        cd = dr.get_ade_coefficients(dt)
        assert np.array_equal([cd["C0"], cd["C1"]], GOLD[f"D_drude_ade_{key}"])
        cb = db.get_ade_coefficients(dt)
        assert np.array_equal([cb["C0"], cb["C1"]], GOLD[f"D_debye_ade_{key}"])


def test_c1_time_step_closed_form():
    """dt = 0.99 / (c sqrt(1/dx^2 + 1/dy^2)) for the 2-D strip at 20 nm."""
    g = _grid("C1exact")
    assert g.get_time_step(0.99) == pytest.approx(4.6701355867643756e-17, rel=1e-15)
    assert float(GOLD["C1exact_dt"]) == pytest.approx(4.6701355867643756e-17, rel=1e-15)
