"""Write golden vectors from the reference's own modules.

No test of the reference holds field, spectrum or coefficient values.  The
reference package cannot be imported whole on Linux: its
``src/prismo/backends/metal_backend.py:85`` evaluates the annotation
``MTLBuffer`` at class creation, which is undefined when pyobjc's Metal is
absent, so ``import prismo`` raises NameError.  The modules with no package
imports -- ``core/grid.py``, ``sources/waveform.py``,
``materials/dispersion.py`` -- load on their own from their files, and this
script runs them on seeded inputs and stores the outputs as fixtures
(``reference_vectors.npz``).  Run in the build container only:

    python tests/golden/make_golden.py
"""

from __future__ import annotations

import importlib.util
from pathlib import Path

import numpy as np

REF = Path("/root/reference/src/prismo")
OUT = Path(__file__).resolve().parent / "reference_vectors.npz"
SEED = 20261202


def load(name, rel):
    spec = importlib.util.spec_from_file_location(name, REF / rel)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def region(ix):
    """(first index, count) per axis of an np.ix_ region; first = -1 if empty."""
    out = []
    for a in range(3):
        if a >= len(ix):
            out += [0, 1]
            continue
        v = ix[a].ravel()
        out += [int(v[0]) if v.size else -1, int(v.size)]
    return np.array(out)


def main():
    grid = load("ref_grid", "core/grid.py")
    wave = load("ref_waveform", "sources/waveform.py")
    disp = load("ref_dispersion", "materials/dispersion.py")
    rng = np.random.default_rng(SEED)
    out = {}

    # ---- grids: dimensions, dt, Courant, point_to_index, component regions
    specs = {
        "C1": ((10e-6, 5e-6, 0.0), 50e6, 20, 0.99),
        "C1exact": ((500 * 20e-9 * (1 - 1e-9), 250 * 20e-9 * (1 - 1e-9), 0.0), 50e6, 20, 0.99),
        "C2": ((232 * 20e-9 * (1 - 1e-9),) * 3, 50e6, 12, 0.99),
        "T3": ((10.0, 5.0, 3.0), 20.0, 5, 0.9),
        "T2": ((8.0, 6.0, 0.0), 25.0, 8, 0.5),
    }
    comps = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
    for tag, (size, res, pml, cf) in specs.items():
        g = grid.YeeGrid(grid.GridSpec(size=size, resolution=res, boundary_layers=pml))
        out[f"{tag}_dims"] = np.array(g.dimensions)
        out[f"{tag}_phys"] = np.array(g.physical_dimensions)
        out[f"{tag}_dt"] = np.array(g.get_time_step(cf))
        out[f"{tag}_courant"] = np.array(g.get_courant_number(g.get_time_step(cf)))
        out[f"{tag}_origin"] = np.asarray(g.origin)
        for c in comps:
            out[f"{tag}_shape_{c}"] = np.array(g.get_field_shape(c))
        # random physical points -> indices, and random boxes -> regions
        lo = g.origin
        span = np.array([g.Nx_total * g.dx, g.Ny_total * g.dy, g.Nz_total * g.dz])
        pts = lo + rng.uniform(-0.05, 1.05, size=(40, 3)) * span
        out[f"{tag}_pts"] = pts
        out[f"{tag}_pidx"] = np.array([g.point_to_index(tuple(p)) for p in pts])
        boxes = rng.integers(-3, max(g.dimensions) + 3, size=(12, 6))
        out[f"{tag}_boxes"] = boxes
        for c in comps:
            out[f"{tag}_reg_{c}"] = np.array(
                [region(g.get_component_indices(c, *map(int, b))) for b in boxes])
    # the C1 line source / monitor regions as the configuration places them
    g1 = grid.YeeGrid(grid.GridSpec(size=specs["C1exact"][0], resolution=50e6, boundary_layers=20))
    for tag, cx in (("src", 1e-6), ("mon", 8e-6)):
        x0, y0, z0 = g1.point_to_index((cx, 2.5e-6 - 1e-6, 0.0))
        x1, y1, z1 = g1.point_to_index((cx, 2.5e-6 + 1e-6, 0.0))
        if x0 == x1:
            x1 = x0 + 1
        for c in ("Ez", "Hy"):
            out[f"C1_{tag}_{c}"] = region(g1.get_component_indices(c, x0, x1, y0, y1, z0, z1))

    # ---- waveforms at seeded times
    ts = np.sort(rng.uniform(0, 100e-15, 64))
    out["W_t"] = ts
    cases = {
        "gauss": wave.GaussianPulse(193e12, 10e-15),
        "gauss_ph": wave.GaussianPulse(150e12, 7e-15, 2.0, 0.4, 20e-15),
        "cw": wave.ContinuousWave(193e12, 1.5, 0.2),
        "ricker": wave.RickerWavelet(100e12, 1.2),
    }
    for k, w in cases.items():
        out[f"W_{k}"] = np.array([w(float(t)) for t in ts])

    # ---- dispersion: closed-form permittivity and ADE coefficients
    w = np.linspace(0.5e15, 3e16, 50)
    si = disp.LorentzMaterial(7.98, [disp.LorentzPole(omega_0=1.1e16, delta_epsilon=3.69,
                                                      gamma=1e13)])
    two = disp.LorentzMaterial(2.0, [disp.LorentzPole(2e15, 1.5, 3e13),
                                     disp.LorentzPole(9e15, 0.7, 1e14)])
    dr = disp.DrudeMaterial(1.0, 1.37e16, 1.07e14)
    db = disp.DebyeMaterial(2.0, 5.0, 1e-14)
    out["D_w"] = w
    out["D_si_eps"] = si.permittivity(w)
    out["D_two_eps"] = two.permittivity(w)
    out["D_drude_eps"] = dr.permittivity(w)
    out["D_debye_eps"] = db.permittivity(w)
    for dt in (3.8131497e-17, 4.67e-17, 1e-16):
        key = f"{dt:.3e}"
        for name, m in (("si", si), ("two", two)):
            co = m.get_ade_coefficients(dt)
            out[f"D_{name}_ade_{key}"] = np.array(
                [[p["C0"], p["C1"], p["C2"], p["C3"]] for p in co["poles"]])
        cd = dr.get_ade_coefficients(dt)
        out[f"D_drude_ade_{key}"] = np.array([cd["C0"], cd["C1"]])
        cb = db.get_ade_coefficients(dt)
        out[f"D_debye_ade_{key}"] = np.array([cb["C0"], cb["C1"]])

    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()
