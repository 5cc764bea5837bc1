"""O0 -- CPU restatement of the reference's time-step arithmetic.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, to check and time; never by the product.

Follows, operation for operation:
  src/prismo/core/solver.py:113-133   Ca/Cb/Da/Db           (coefficients)
  src/prismo/core/solver.py:167-253   3-D H update          (h3d)
  src/prismo/core/solver.py:255-309   3-D E update          (e3d)
  src/prismo/core/solver.py:311-397   2-D H update + gates  (h2d)
  src/prismo/core/solver.py:399-456   2-D E update          (e2d)
  src/prismo/core/solver.py:458-533   averaging, redone every call as there
  src/prismo/core/simulation.py:147-164  step order, FP64 time accumulation
  src/prismo/monitors/field.py:128-143   running DFT of a region
Field arrays use the reference shapes and C order.
"""

from __future__ import annotations

import numpy as np

EPS0 = 8.854187817e-12
MU0 = 4 * np.pi * 1e-7


def coefficients(eps_rel, mu_rel, sigma_e, sigma_m, dt):
    """solver.py:119-130."""
    eps = EPS0 * eps_rel
    s = sigma_e * dt / (2 * eps)
    Ca = (1 - s) / (1 + s)
    Cb = (dt / eps) / (1 + s)
    mu = MU0 * mu_rel
    sm = sigma_m * dt / (2 * mu)
    Da = (1 - sm) / (1 + sm)
    Db = (dt / mu) / (1 + sm)
    return Ca, Cb, Da, Db


def shapes(nx, ny, nz, is_2d):
    """grid.py:137-176."""
    if is_2d:
        return {"Ex": (nx, ny - 1), "Ey": (nx - 1, ny), "Ez": (nx - 1, ny - 1),
                "Hx": (nx - 1, ny), "Hy": (nx, ny - 1), "Hz": (nx, ny)}
    return {"Ex": (nx, ny - 1, nz - 1), "Ey": (nx - 1, ny, nz - 1), "Ez": (nx - 1, ny - 1, nz),
            "Hx": (nx - 1, ny, nz), "Hy": (nx, ny - 1, nz), "Hz": (nx, ny, nz - 1)}


def zeros(nx, ny, nz, is_2d, dtype=np.float64):
    return {k: np.zeros(v, dtype=dtype) for k, v in shapes(nx, ny, nz, is_2d).items()}


# -- averaging (solver.py:458-533) ---------------------------------------------
def _ex(a): return 0.25 * (a[:, :-1, :-1] + a[:, 1:, :-1] + a[:, :-1, 1:] + a[:, 1:, 1:])
def _ey(a): return 0.25 * (a[:-1, :, :-1] + a[1:, :, :-1] + a[:-1, :, 1:] + a[1:, :, 1:])
def _ez(a): return 0.25 * (a[:-1, :-1, :] + a[1:, :-1, :] + a[:-1, 1:, :] + a[1:, 1:, :])
def _hx(a): return 0.5 * (a[:-1, :, :] + a[1:, :, :])
def _hy(a): return 0.5 * (a[:, :-1, :] + a[:, 1:, :])
def _hz(a): return 0.5 * (a[:, :, :-1] + a[:, :, 1:])
def _ex2(a): return 0.5 * (a[:, :-1] + a[:, 1:])
def _ey2(a): return 0.5 * (a[:-1, :] + a[1:, :])
def _ez2(a): return 0.25 * (a[:-1, :-1] + a[1:, :-1] + a[:-1, 1:] + a[1:, 1:])
def _hx2(a): return 0.5 * (a[:-1, :] + a[1:, :])
def _hy2(a): return 0.5 * (a[:, :-1] + a[:, 1:])


def h3d(F, C, d):
    Ex, Ey, Ez, Hx, Hy, Hz = (F[k] for k in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"))
    Da, Db = C["Da"], C["Db"]
    dx, dy, dz = d
    cy = (Ez[:, 1:, :] - Ez[:, :-1, :]) / dy
    cz = (Ey[:, :, 1:] - Ey[:, :, :-1]) / dz
    nyc = min(Hx.shape[1], cy.shape[1], cz.shape[1])
    nzc = min(Hx.shape[2], cy.shape[2], cz.shape[2])
    da = _hx(Da)[: Hx.shape[0], :nyc, :nzc]
    db = _hx(Db)[: Hx.shape[0], :nyc, :nzc]
    Hx[:, :nyc, :nzc] = da * Hx[:, :nyc, :nzc] - db * (cy[:, :nyc, :nzc] - cz[:, :nyc, :nzc])

    cz = (Ex[:, :, 1:] - Ex[:, :, :-1]) / dz
    cx = (Ez[1:, :, :] - Ez[:-1, :, :]) / dx
    nxc = min(Hy.shape[0], cz.shape[0], cx.shape[0])
    nzc = min(Hy.shape[2], cz.shape[2], cx.shape[2])
    da = _hy(Da)[:nxc, : Hy.shape[1], :nzc]
    db = _hy(Db)[:nxc, : Hy.shape[1], :nzc]
    Hy[:nxc, :, :nzc] = da * Hy[:nxc, :, :nzc] - db * (cz[:nxc, :, :nzc] - cx[:nxc, :, :nzc])

    cx = (Ey[1:, :, :] - Ey[:-1, :, :]) / dx
    cy = (Ex[:, 1:, :] - Ex[:, :-1, :]) / dy
    nxc = min(Hz.shape[0], cx.shape[0], cy.shape[0])
    nyc = min(Hz.shape[1], cx.shape[1], cy.shape[1])
    da = _hz(Da)[:nxc, :nyc, : Hz.shape[2]]
    db = _hz(Db)[:nxc, :nyc, : Hz.shape[2]]
    Hz[:nxc, :nyc, :] = da * Hz[:nxc, :nyc, :] - db * (cx[:nxc, :nyc, :] - cy[:nxc, :nyc, :])


def e3d(F, C, d):
    Ex, Ey, Ez, Hx, Hy, Hz = (F[k] for k in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"))
    Ca, Cb = C["Ca"], C["Cb"]
    dx, dy, dz = d
    c1 = (Hz[:, 1:, :] - Hz[:, :-1, :]) / dy
    c2 = (Hy[:, :, 1:] - Hy[:, :, :-1]) / dz
    ca = _ex(Ca)[: Ex.shape[0], : Ex.shape[1], : Ex.shape[2]]
    cb = _ex(Cb)[: Ex.shape[0], : Ex.shape[1], : Ex.shape[2]]
    Ex[:, :, :] = ca * Ex + cb * (c1 - c2)
    c1 = (Hx[:, :, 1:] - Hx[:, :, :-1]) / dz
    c2 = (Hz[1:, :, :] - Hz[:-1, :, :]) / dx
    ca = _ey(Ca)[: Ey.shape[0], : Ey.shape[1], : Ey.shape[2]]
    cb = _ey(Cb)[: Ey.shape[0], : Ey.shape[1], : Ey.shape[2]]
    Ey[:, :, :] = ca * Ey + cb * (c1 - c2)
    c1 = (Hy[1:, :, :] - Hy[:-1, :, :]) / dx
    c2 = (Hx[:, 1:, :] - Hx[:, :-1, :]) / dy
    ca = _ez(Ca)[: Ez.shape[0], : Ez.shape[1], : Ez.shape[2]]
    cb = _ez(Cb)[: Ez.shape[0], : Ez.shape[1], : Ez.shape[2]]
    Ez[:, :, :] = ca * Ez + cb * (c1 - c2)


def h2d(F, C, d):
    Ez, Hx, Hy = F["Ez"], F["Hx"], F["Hy"]
    Da, Db = C["Da"][:, :, 0], C["Db"][:, :, 0]
    dx, dy = d[0], d[1]
    if np.max(np.abs(Ez)) > 0:
        nxh, nyh = Hx.shape
        nxe, nye = Ez.shape
        if nxe >= nxh and nye > 0:
            cy = (Ez[:nxh, 1:] - Ez[:nxh, :-1]) / dy
            da, db = _hx2(Da), _hx2(Db)
            n = min(nyh, cy.shape[1])
            Hx[:, :n] = da[:nxh, :n] * Hx[:, :n] + db[:nxh, :n] * cy[:, :n]
        nxh, nyh = Hy.shape
        if nxe > 0 and nye >= nyh:
            cx = (Ez[1:, :nyh] - Ez[:-1, :nyh]) / dx
            da, db = _hy2(Da), _hy2(Db)
            n = min(nxh, cx.shape[0])
            Hy[:n, :] = da[:n, :nyh] * Hy[:n, :] - db[:n, :nyh] * cx[:n, :]
    Hz, Ex, Ey = F["Hz"], F["Ex"], F["Ey"]
    if np.max(np.abs(Ex)) > 0 or np.max(np.abs(Ey)) > 0:
        cx = np.zeros_like(Hz)
        if Ey.shape[0] > 1:
            n = min(Ey.shape[0] - 1, Hz.shape[0] - 1)
            cx[:n, :] = (Ey[1:n + 1, :] - Ey[:n, :]) / dx
        cy = np.zeros_like(Hz)
        if Ex.shape[1] > 1:
            n = min(Ex.shape[1] - 1, Hz.shape[1] - 1)
            cy[:, :n] = (Ex[:, 1:n + 1] - Ex[:, :n]) / dy
        nx, ny = Hz.shape
        Hz[:, :] = Da[:nx, :ny] * Hz - Db[:nx, :ny] * (cx - cy)


def e2d(F, C, d):
    Ez, Hx, Hy = F["Ez"], F["Hx"], F["Hy"]
    Ca, Cb = C["Ca"][:, :, 0], C["Cb"][:, :, 0]
    dx, dy = d[0], d[1]
    if Ez.shape[0] > 0 and Ez.shape[1] > 0:
        c1 = (Hy[1:, :] - Hy[:-1, :]) / dx
        c2 = (Hx[:, 1:] - Hx[:, :-1]) / dy
        nx, ny = Ez.shape
        ca = _ez2(Ca)[:nx, :ny]
        cb = _ez2(Cb)[:nx, :ny]
        Ez[:, :] = ca * Ez + cb * (c1[:nx, :ny] - c2[:nx, :ny])
    Hz, Ex, Ey = F["Hz"], F["Ex"], F["Ey"]
    if Hz.shape[0] > 0 and Hz.shape[1] > 0:
        if Ex.shape[0] > 0 and Ex.shape[1] > 0:
            c = (Hz[:, 1:] - Hz[:, :-1]) / dy
            nx, ny = Ex.shape
            Ex[:, :] = _ex2(Ca)[:nx, :ny] * Ex + _ex2(Cb)[:nx, :ny] * c
        if Ey.shape[0] > 0 and Ey.shape[1] > 0:
            c = (Hz[1:, :] - Hz[:-1, :]) / dx
            nx, ny = Ey.shape
            Ey[:, :] = _ey2(Ca)[:nx, :ny] * Ey - _ey2(Cb)[:nx, :ny] * c


def step(F, C, d, is_2d):
    """MaxwellUpdater.step (solver.py:535-552)."""
    if is_2d:
        h2d(F, C, d)
        e2d(F, C, d)
    else:
        h3d(F, C, d)
        e3d(F, C, d)


def run(F, C, d, is_2d, dt, nsteps, t0=0.0, sources=(), monitors=()):
    """Simulation.step loop (simulation.py:147-164).

    sources: callables (F, t) -> None applied in order after each step.
    monitors: objects with .update(F, t, dt).  Returns the final time.
    """
    t = t0
    for _ in range(nsteps):
        step(F, C, d, is_2d)
        t += dt
        for s in sources:
            s(F, t)
        for m in monitors:
            m.update(F, t, dt)
    return t


class RegionDFT:
    """FieldMonitor frequency branch (field.py:128-143) on index regions."""

    def __init__(self, regions: dict, freqs):
        self.regions = regions  # comp -> np.ix_ tuple
        self.freqs = list(freqs)
        self.data = {c: {f: 0 for f in self.freqs} for c in regions}
        self.samples = {c: [] for c in regions}
        self.times = []

    def update(self, F, t, dt):
        self.times.append(t)
        for c, idx in self.regions.items():
            x = F[c][idx].copy()
            self.samples[c].append(x)
            for f in self.freqs:
                omega = 2 * np.pi * f
                self.data[c][f] = self.data[c][f] + x * np.exp(-1j * omega * t) * dt
