"""Device rasterisation and coefficient build (include/pmgeom.h) against the
shape oracle (oracle/geometry.py) and the host coefficient path."""

import numpy as np
import pytest

from oracle import geometry as G

pytestmark = pytest.mark.gpu

SEED = 20261202


def _random_shapes(rng, n, lo, hi, two_d=False):
    """Oracle dicts and the equivalent prismo_amd shapes."""
    import prismo_amd as pa

    mats = [pa.Material(f"m{k}", float(k + 2)) for k in range(4)]
    od, ps = [], []
    for t in range(n):
        kind = ["box", "sphere", "cylinder", "polygon"][t % 4]
        c = rng.uniform(lo, hi, 3)
        if two_d:
            c[2] = 0.0
        m = int(rng.integers(0, 4))
        if kind == "box":
            size = rng.uniform(0.1, 0.6, 3) * (hi - lo)
            if two_d:
                size[2] = 0.0
            od.append({"kind": "box", "center": c, "size": size, "material": m + 1})
            ps.append(pa.Box(mats[m], c, size))
        elif kind == "sphere":
            r = float(rng.uniform(0.1, 0.4) * (hi - lo))
            od.append({"kind": "sphere", "center": c, "radius": r, "material": m + 1})
            ps.append(pa.Sphere(mats[m], c, r))
        elif kind == "cylinder":
            ax = "xyz"[int(rng.integers(0, 3))] if not two_d else "z"
            r = float(rng.uniform(0.1, 0.3) * (hi - lo))
            h = float(rng.uniform(0.2, 0.8) * (hi - lo))
            od.append({"kind": "cylinder", "center": c, "radius": r, "height": h, "axis": ax,
                       "material": m + 1})
            ps.append(pa.Cylinder(mats[m], c, r, h, ax))
        else:
            k = int(rng.integers(3, 8))
            ang = np.sort(rng.uniform(0, 2 * np.pi, k))
            rad = rng.uniform(0.1, 0.4, k) * (hi - lo)
            v = np.column_stack([c[0] + rad * np.cos(ang), c[1] + rad * np.sin(ang)])
            z0, z1 = sorted(rng.uniform(lo, hi, 2))
            if two_d:
                z0, z1 = -np.inf, np.inf
            od.append({"kind": "polygon", "vertices": v, "z_min": z0, "z_max": z1,
                       "material": m + 1})
            ps.append(pa.Polygon(mats[m], v, z0, z1))
    return od, ps, mats


@pytest.mark.parametrize("two_d", [False, True])
def test_raster_matches_oracle(mi355x, two_d):
    import torch

    from prismo_amd.geometry import raster

    rng = np.random.default_rng(SEED)
    x = np.linspace(-1.0, 1.0, 37)
    y = np.linspace(-0.9, 1.1, 29)
    z = None if two_d else np.linspace(-1.2, 0.8, 23)
    od, ps, mats = _random_shapes(rng, 16, -1.0, 1.0, two_d)
    want = G.paint(od, x, y, z)                                  # (nx, ny, nz)
    table = [None] + mats
    idx, _ = raster.paint(ps, x, y, np.zeros(1) if two_d else z, torch.device("cuda", 0),
                          table)
    got = idx.cpu().numpy().transpose(2, 1, 0)                   # -> (nx, ny, nz)
    # table order: material k of the shapes has index k + 1 in both
    assert np.array_equal(got, want)
    # the product's host rasterize gives the same masks
    for o, s in zip(od, ps):
        m = s.rasterize(x, y, z)
        ref = G.paint([dict(o, material=1)], x, y, z)
        assert np.array_equal(m.reshape(ref.shape), ref == 1)


def test_raster_boundary_nodes(mi355x):
    """Nodes exactly on box faces, sphere shells and polygon edges (closed
    box and sphere, half-open ray casting) decide as the oracle does."""
    import torch

    import prismo_amd as pa
    from prismo_amd.geometry import raster

    v = np.arange(-6.0, 7.0)
    m = pa.Material("a", 4.0)
    shapes = [pa.Box(m, (0.0, 0.0, 0.0), (4.0, 6.0, 2.0)), pa.Sphere(m, (1.0, -1.0, 0.0), 3.0),
              pa.Polygon(m, [(-5.0, -5.0), (5.0, -5.0), (0.0, 5.0)], -2.0, 2.0),
              pa.Cylinder(m, (0.0, 2.0, 0.0), 2.0, 4.0, "x")]
    od = [{"kind": "box", "center": (0.0, 0.0, 0.0), "size": (4.0, 6.0, 2.0)},
          {"kind": "sphere", "center": (1.0, -1.0, 0.0), "radius": 3.0},
          {"kind": "polygon", "vertices": [(-5.0, -5.0), (5.0, -5.0), (0.0, 5.0)],
           "z_min": -2.0, "z_max": 2.0},
          {"kind": "cylinder", "center": (0.0, 2.0, 0.0), "radius": 2.0, "height": 4.0,
           "axis": "x"}]
    for s, o in zip(shapes, od):
        idx, _ = raster.paint([s], v, v, v, torch.device("cuda", 0), [None, m])
        got = idx.cpu().numpy().transpose(2, 1, 0)
        assert np.array_equal(got, G.paint([dict(o, material=1)], v, v, v)), o["kind"]


def test_cb_fp16_rounding_matches_numpy(mi355x):
    """FP16 Cb / scale rounds once from FP64 to nearest even, as numpy's
    float64 -> float16 conversion (normal, subnormal, overflow, ties)."""
    import torch

    from prismo_amd import _native as N
    from prismo_amd.geometry import raster

    rng = np.random.default_rng(SEED + 1)
    scale = 1.0
    lut = np.concatenate([
        rng.uniform(0.0, 1.0, 40), 10.0 ** rng.uniform(-9, 4.9, 80),
        [6.0e-8, 5.96e-8, 2.0 ** -25, 2.0 ** -24, 3 * 2.0 ** -26, 65504.0, 65519.0, 65520.0,
         1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, 0.0],
    ])[:255]
    n_lut = len(lut)
    nx, ny, nz = 9, 8, 7
    dev = torch.device("cuda", 0)
    idx = torch.from_numpy(rng.integers(0, n_lut, (nz, ny, nx)).astype(np.uint8)).to(dev)
    sy, sz = 16, 16 * ny
    for comp in range(3):
        out = torch.zeros((nz, ny, sy), dtype=torch.float16, device=dev)
        raster.cb_from_index(idx, lut, comp, N.CB_F16, scale, out, sy, sz)
        node = lut[idx.cpu().numpy().transpose(2, 1, 0)]        # (nx, ny, nz)
        from prismo_amd.core.coefficients import component_average

        avg = component_average(node, ("Ex", "Ey", "Ez")[comp], False)
        want = (avg / scale).astype(np.float16)
        got = out.cpu().numpy()[: avg.shape[2], : avg.shape[1], : avg.shape[0]].transpose(2, 1, 0)
        assert np.array_equal(got.view(np.uint16), want.view(np.uint16)), comp


@pytest.mark.parametrize("dtype,storage", [(np.float64, "auto"), (np.float32, "fp16")])
def test_structures_equal_material_arrays(mi355x, dtype, storage):
    """A simulation painted from shapes on the device steps bit for bit as
    the same structure given as host node arrays."""
    import prismo_amd as pa

    D = 20e-9
    n = (40, 24, 20)
    size = tuple((k - 8) * D * (1 - 1e-9) for k in n)
    si, ox = pa.Material("Si", 12.1), pa.Material("SiO2", 2.085)
    shapes = [pa.Box(ox, (0.0, 0.0, -0.1e-6), (2e-6, 2e-6, 0.2e-6)),
              pa.Box(si, (0.0, 0.0, 0.05e-6), (2e-6, 0.12e-6, 0.1e-6)),
              pa.Cylinder(si, (0.1e-6, 0.1e-6, 0.0), 0.06e-6, 0.2e-6, "z")]

    def build(host):
        kw = dict(size=size, resolution=1 / D, pml_layers=4, courant_factor=0.99,
                  scheme="yee", dtype=dtype, pml=pa.PMLParams(4), coefficient_storage=storage)
        if host:
            g = pa.Simulation(**kw).grid
            nx, ny, nz = g.dimensions
            x = g.origin[0] + np.arange(nx) * g.dx
            y = g.origin[1] + np.arange(ny) * g.dy
            z = g.origin[2] + np.arange(nz) * g.dz
            eps = np.full((nx, ny, nz), 1.5)
            for s in shapes:
                eps[s.rasterize(x, y, z)] = s.material.epsilon_r
            sim = pa.Simulation(material_arrays={"eps_rel": eps}, **kw)
        else:
            sim = pa.Simulation(**kw)
            sim.set_background(1.5)
            for s in shapes:
                sim.add_structure(s)
        sim.add_source(pa.ElectricDipole(position=(0.0, 0.0, 0.05e-6), polarization="y",
                                         frequency=193e12, pulse_width=5e-15))
        sim.run(40 * sim.dt * 0.9999)
        return {c: sim.fields.to_numpy(c) for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")}

    a, b = build(False), build(True)
    for c in a:
        assert np.array_equal(a[c], b[c]), c
    assert np.abs(a["Ey"]).max() > 0
