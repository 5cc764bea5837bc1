"""O1 -- CPU textbook Yee oracle with CPML, coupled ADE and region DFT.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, to check and time; never by the product.

Scheme (stated here, pinned by the physics tests in tests/test_oracle_*):
  * staggering of src/prismo/core/solver.py:173-175,261-263 (E on faces, H on
    edges); E update = solver.py:255-309 (forward differences of H);
  * H update with BACKWARD differences of E, E beyond the grid = 0 (PEC wall
    half a cell outside the outer H nodes) -- the consistent Yee partner of
    that E update;
  * 2-D TM (Ez, Hx, Hy): Hx -= Db dEz/dy, Hy += Db dEz/dx, Ez += Cb curl H;
  * coefficients averaged to component points as solver.py:458-533;
  * CPML (Roden & Gedney 2000), full-size psi arrays, b/c/kappa profiles from
    ``cpml_profiles`` below (an independent restatement of the textbook
    grading, compared with the product's in the tests);
  * ADE with the coefficients of src/prismo/materials/dispersion.py:189-231
    (Lorentz C0..C3) and 267-286 (Drude C0, C1), coupled through
    D = eps_inf E + P (units of eps0) with the drive averaged over the two
    half steps around each update (time-centred):
      Lorentz  P+ = C2 P + C3 P- + [(C0 E+ + C1 E) + (C0 E + C1 E-)] / 2
      Drude    J+ = C1 J + C0 (E+ + E) / 2,   P+ - P = dt (J+ + J) / 2;
  * sources add after the E update, monitors sample after sources
    (src/prismo/core/simulation.py:147-164).
"""

from __future__ import annotations

import numpy as np

EPS0 = 8.854187817e-12
MU0 = 4 * np.pi * 1e-7
ETA0 = np.sqrt(MU0 / EPS0)


def shapes(nx, ny, nz, is_2d):
    if is_2d:
        return {"Ez": (nx - 1, ny - 1), "Hx": (nx - 1, ny), "Hy": (nx, ny - 1)}
    return {"Ex": (nx, ny - 1, nz - 1), "Ey": (nx - 1, ny, nz - 1), "Ez": (nx - 1, ny - 1, nz),
            "Hx": (nx - 1, ny, nz), "Hy": (nx, ny - 1, nz), "Hz": (nx, ny, nz - 1)}


def zeros(nx, ny, nz, is_2d, dtype=np.float64):
    return {k: np.zeros(v, dtype=dtype) for k, v in shapes(nx, ny, nz, is_2d).items()}


def bdiff(a, axis):
    """Backward difference with zero beyond both ends: out[j] = a[j] - a[j-1],
    j = 0..n (length n+1 along axis)."""
    pad = [(0, 0)] * a.ndim
    pad[axis] = (1, 1)
    return np.diff(np.pad(a, pad), axis=axis)


def cpml_profiles(n, t_lo, t_hi, d, dt, m=3, R=1e-8, kappa_max=15.0, alpha_max=0.0,
                  sigma_max=None):
    """b, c, 1/kappa along one axis, at half nodes ('e') and nodes ('h').

    Walls at positions -1/2 and n-1/2; rho = depth / t inside a layer."""
    out = {}
    for tag, off in (("e", 0.5), ("h", 0.0)):
        p = np.arange(n) + off
        b = np.ones(n)
        c = np.zeros(n)
        ik = np.ones(n)
        for t, lo in ((t_lo, True), (t_hi, False)):
            if t <= 0:
                continue
            smax = sigma_max if sigma_max is not None else -(m + 1) * np.log(R) / (2 * ETA0 * t * d)
            sel = np.arange(n) < t if lo else np.arange(n) >= n - t
            rho = ((t - 0.5 - p) / t) if lo else ((p - (n - 0.5 - t)) / t)
            rho = np.clip(rho, 0, 1)[sel]
            sg = smax * rho**m
            ka = 1 + (kappa_max - 1) * rho**m
            al = alpha_max * (1 - rho) ** m
            bb = np.exp(-(sg / ka + al) * dt / EPS0)
            den = sg * ka + ka * ka * al
            with np.errstate(invalid="ignore", divide="ignore"):
                cc = np.where(den > 0, sg / den * (bb - 1), 0.0)
            b[sel], c[sel], ik[sel] = bb, cc, 1 / ka
        out["b" + tag], out["c" + tag], out["ik" + tag] = b, c, ik
    return out


def _bc(prof, axis, ndim):
    shape = [1] * ndim
    shape[axis] = -1
    return prof.reshape(shape)


class YeeOracle:
    """State + one-step update.  F: dict of field arrays (shapes above);
    coef: dict comp -> coefficient array at that component's points
    ("cb" for E comps, optionally "ca"; H uses "da"/"db" arrays or a scalar
    db with da = 1)."""

    def __init__(self, nx, ny, nz, d, dt, cb, db, ca=None, da=None, pml=None,
                 pml_kw=None, dtype=np.float64):
        self.is_2d = nz == 1
        self.n = (nx, ny, nz)
        self.d = d
        self.dt = dt
        self.dtype = dtype
        self.F = zeros(nx, ny, nz, self.is_2d, dtype)
        self.cb, self.db, self.ca, self.da = cb, db, ca or {}, da or {}
        self.pml = pml  # (x-, x+, y-, y+, z-, z+) or None
        self.prof = None
        self.psi = {}
        self.ade = []
        if pml is not None:
            kw = pml_kw or {}
            self.prof = []
            for a in range(2 if self.is_2d else 3):
                tl, th = pml[2 * a], pml[2 * a + 1]
                self.prof.append(cpml_profiles(self.n[a], tl, th, d[a], dt, **kw)
                                 if tl + th else None)

    # ---------------------------------------------------------------- helpers
    def _term(self, key, d_arr, axis, side):
        """CPML-stretched derivative term (side 'e' or 'h'):
        psi <- b psi + c d, term = d / kappa + psi.  Outside the layers
        b = 1, c = 0, 1/kappa = 1 and psi stays 0, so the term is d itself;
        only the layer slabs along ``axis`` are computed (same values as the
        full-array form)."""
        if self.prof is None or axis >= len(self.prof) or self.prof[axis] is None:
            return d_arr
        pr = self.prof[axis]
        nd = d_arr.ndim
        L = d_arr.shape[axis]
        tl, th = self.pml[2 * axis], self.pml[2 * axis + 1]
        n = self.n[axis]
        slabs = [s for s in ((0, min(tl, L)), (max(n - th, 0), L)) if s[1] > s[0]]
        psis = self.psi.get(key)
        if psis is None:
            psis = []
            for a0, a1 in slabs:
                shp = list(d_arr.shape)
                shp[axis] = a1 - a0
                psis.append(np.zeros(shp, dtype=d_arr.dtype))
            self.psi[key] = psis
        out = d_arr.copy()
        for (a0, a1), psi in zip(slabs, psis):
            sl = [slice(None)] * nd
            sl[axis] = slice(a0, a1)
            sl = tuple(sl)
            b = _bc(pr["b" + side][a0:a1], axis, nd)
            c = _bc(pr["c" + side][a0:a1], axis, nd)
            ik = _bc(pr["ik" + side][a0:a1], axis, nd)
            d = d_arr[sl]
            psi[...] = b * psi + c * d
            out[sl] = d * ik + psi
        return out

    def add_ade(self, comp, box, kind, eps_inf, c):
        (i0, i1), (j0, j1), (k0, k1) = box
        shp = (i1 - i0, j1 - j0) if self.is_2d else (i1 - i0, j1 - j0, k1 - k0)
        self.ade.append(dict(comp=comp, box=box, kind=kind, eps_inf=eps_inf, c=c,
                             P=np.zeros(shp, self.dtype), Pm=np.zeros(shp, self.dtype),
                             Em=np.zeros(shp, self.dtype)))

    def _sl(self, box):
        (i0, i1), (j0, j1), (k0, k1) = box
        return (slice(i0, i1), slice(j0, j1)) if self.is_2d else (
            slice(i0, i1), slice(j0, j1), slice(k0, k1))

    # ---------------------------------------------------------------- update
    def update_h(self):
        F = self.F
        dx, dy, dz = self.d
        db = self.db
        if self.is_2d:
            Ez = F["Ez"]
            d1 = bdiff(Ez, 1)[: F["Hx"].shape[0]] / dy          # dEz/dy at Hx
            d1 = self._term("hx1", d1, 1, "h")
            self._hset("Hx", d1)
            d2 = bdiff(Ez, 0)[:, : F["Hy"].shape[1]] / dx       # dEz/dx at Hy
            d2 = self._term("hy2", d2, 0, "h")
            self._hset("Hy", -d2)
            return
        Ex, Ey, Ez = F["Ex"], F["Ey"], F["Ez"]
        self._hset("Hx", self._term("hx1", bdiff(Ez, 1) / dy, 1, "h")
                   - self._term("hx2", bdiff(Ey, 2) / dz, 2, "h"))
        self._hset("Hy", self._term("hy1", bdiff(Ex, 2) / dz, 2, "h")
                   - self._term("hy2", bdiff(Ez, 0) / dx, 0, "h"))
        self._hset("Hz", self._term("hz1", bdiff(Ey, 0) / dx, 0, "h")
                   - self._term("hz2", bdiff(Ex, 1) / dy, 1, "h"))

    def _hset(self, c, curl):
        H = self.F[c]
        da = self.da.get(c, 1.0)
        db = self.db[c] if isinstance(self.db, dict) else self.db
        H[...] = da * H - db * curl

    def update_e(self):
        F = self.F
        dx, dy, dz = self.d
        if self.is_2d:
            Hx, Hy = F["Hx"], F["Hy"]
            c = (self._term("ez1", np.diff(Hy, axis=0) / dx, 0, "e")
                 - self._term("ez2", np.diff(Hx, axis=1) / dy, 1, "e"))
            self._eset("Ez", c)
            return
        Hx, Hy, Hz = F["Hx"], F["Hy"], F["Hz"]
        self._eset("Ex", self._term("ex1", np.diff(Hz, axis=1) / dy, 1, "e")
                   - self._term("ex2", np.diff(Hy, axis=2) / dz, 2, "e"))
        self._eset("Ey", self._term("ey1", np.diff(Hx, axis=2) / dz, 2, "e")
                   - self._term("ey2", np.diff(Hz, axis=0) / dx, 0, "e"))
        self._eset("Ez", self._term("ez1", np.diff(Hy, axis=0) / dx, 0, "e")
                   - self._term("ez2", np.diff(Hx, axis=1) / dy, 1, "e"))

    def _eset(self, c, curl):
        E = self.F[c]
        new = self.ca.get(c, 1.0) * E + self.cb[c] * curl
        for r in self.ade:
            if r["comp"] != c:
                continue
            sl = self._sl(r["box"])
            e = E[sl]
            s = (self.dt / EPS0) * curl[sl]
            einf = r["eps_inf"]
            c0, c1, c2, c3 = r["c"]
            if r["kind"] == 1:
                # keep P-, E- and solve D+ = D + s with
                # P+ = C2 P + C3 P- + (C0 E+ + (C0 + C1) E + C1 E-) / 2
                P, Pm, Em = r["P"], r["Pm"], r["Em"]
                known = c2 * P + c3 * Pm + 0.5 * ((c0 + c1) * e + c1 * Em)
                en = (einf * e + P + s - known) / (einf + 0.5 * c0)
                Pn = known + 0.5 * c0 * en
                r["Pm"], r["P"], r["Em"] = P, Pn, e.copy()
            else:
                J = r["P"]
                h = 0.5 * self.dt
                # eps_inf (E+ - E) + h (J+ + J) = s,  J+ = C1 J + C0 (E+ + E) / 2
                en = (einf * e + s - h * J - h * (c1 * J + 0.5 * c0 * e)) / (einf + 0.5 * h * c0)
                r["P"] = c1 * J + 0.5 * c0 * (en + e)
            new[sl] = en
        E[...] = new

    def step(self):
        self.update_h()
        self.update_e()

    def run(self, nsteps, dt=None, t0=0.0, sources=(), monitors=()):
        """Simulation.step loop (src/prismo/core/simulation.py:147-164): H,
        E, ``t += dt`` accumulated in FP64, every source at the new time,
        then every monitor.  sources: callables (F, t) -> None; monitors:
        objects with .update(F, t, dt).  Returns the final time."""
        dt = self.dt if dt is None else dt
        t = t0
        for _ in range(nsteps):
            self.step()
            t += dt
            for s in sources:
                s(self.F, t)
            for m in monitors:
                m.update(self.F, t, dt)
        return t


def add_box(F, comp, box, value, is_2d):
    (i0, i1), (j0, j1), (k0, k1) = box
    if is_2d:
        F[comp][i0:i1, j0:j1] += value
    else:
        F[comp][i0:i1, j0:j1, k0:k1] += value


def energy(F, d, is_2d):
    """0.5 (eps0 E^2 + mu0 H^2) dV summed over all samples."""
    dV = d[0] * d[1] * (1 if is_2d else d[2])
    e = sum(np.sum(F[c] ** 2) for c in F if c.startswith("E"))
    h = sum(np.sum(F[c] ** 2) for c in F if c.startswith("H"))
    return 0.5 * (EPS0 * e + MU0 * h) * dV
