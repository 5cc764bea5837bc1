"""Builders for the benchmark workloads (synthetic, seeded).

* ``strip_2d``   -- 2-D TM, 500 x 250 physical cells + 20-cell layers at
  20 nm, Courant 0.99 of the 2-D limit, vacuum with a 500 nm Si strip
  (eps 12.1) along x, 2 um line source at x = 1 um (Gaussian 193 THz,
  10 fs), 11-frequency line monitor at x = 8 um (187-200 THz), 3000 steps.
* ``voxel_3d``   -- 256^3 total cells at 20 nm with 12-cell CPML inside
  (232^3 physical + 12 each side), eps per 8^3 block drawn from
  {1.0, 2.085, 12.1} with numpy.random.default_rng(20261202), Ez dipole at
  the centre with a 193 THz Gaussian pulse (10 fs, delay 30 fs), Courant
  0.99 of the 3-D limit, 1000 steps.
"""

from __future__ import annotations

import numpy as np

SEED = 20261202
D = 20e-9
SHRINK = 1 - 1e-9  # avoids ceil() adding a cell to exact products


def strip_2d(scheme="compat", dtype=np.float64, pml=True, steps=3000, backend=None):
    import prismo_amd as pa

    kw = {}
    if scheme == "yee" and pml:
        kw["pml"] = pa.PMLParams(thickness=20, polynomial_order=3, reflection=1e-8)
    sim = pa.Simulation(size=(500 * D * SHRINK, 250 * D * SHRINK, 0.0), resolution=1 / D,
                        pml_layers=20, courant_factor=0.99, scheme=scheme, dtype=dtype,
                        backend=backend, material_arrays=_strip_eps(), **kw)
    sim.add_source(pa.PlaneWaveSource(center=(1e-6, 2.5e-6, 0.0), size=(0.0, 2e-6, 0.0),
                                      direction="+x", polarization="z", frequency=193e12,
                                      pulse_width=10e-15))
    mon = pa.FieldMonitor(center=(8e-6, 2.5e-6, 0.0), size=(0.0, 2e-6, 0.0),
                          components=["Ez", "Hy"], time_domain=False,
                          frequencies=list(np.linspace(187e12, 200e12, 11)))
    sim.add_monitor(mon)
    return sim, mon, steps


def _strip_eps():
    nx, ny = 540, 290
    # node y coordinates, origin at -20 cells
    y = (np.arange(ny) - 20) * D
    eps = np.ones((nx, ny, 1))
    eps[:, np.abs(y - 2.5e-6) <= 250e-9 + 1e-15, 0] = 12.1
    return {"eps_rel": eps}


def voxel_eps(n=256, block=8, seed=SEED):
    rng = np.random.default_rng(seed)
    nb = (n + block - 1) // block
    blocks = rng.choice(np.array([1.0, 2.085, 12.1]), size=(nb, nb, nb))
    return np.repeat(np.repeat(np.repeat(blocks, block, 0), block, 1), block, 2)[:n, :n, :n]


def voxel_3d(dtype=np.float32, coefficient_storage="fp16", n=256, pml=12, backend=None,
             scheme="yee"):
    import prismo_amd as pa

    sim = pa.Simulation(size=((n - 2 * pml) * D * SHRINK,) * 3, resolution=1 / D,
                        pml_layers=pml, courant_factor=0.99, scheme=scheme, dtype=dtype,
                        backend=backend, pml=pa.PMLParams(thickness=pml) if scheme == "yee" else None,
                        material_arrays={"eps_rel": voxel_eps(n)},
                        coefficient_storage=coefficient_storage)
    c = (n // 2 - pml) * D  # physical coordinate of the centre node
    sim.add_source(pa.ElectricDipole(position=(c, c, c), polarization="z", frequency=193e12,
                                     pulse_width=10e-15))
    return sim
