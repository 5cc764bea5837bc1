"""Waveguide eigenmodes on a cross-section (src/prismo/modes/solver.py:16-335).

Same API as the reference (``WaveguideMode``, ``ModeSolver(wavelength, x,
y, epsilon).solve(num_modes, mode_type)``, ``get_mode``), with the solver
itself rebuilt: the reference's operator indexes the flattened (nx, ny)
profile with x-neighbours at +-1 and y-neighbours at +-nx (only right when
nx == ny, and wrapping across rows), targets the middle of the index range
rather than the fundamental, and leaves "TM" and "vector" as empty lists.

Here: semi-vectorial finite differences on the node grid (x, y), zero field
outside it, the dominant transverse E component found from

    TE (E_x dominant):  d/dx[(1/eps) d(eps E_x)/dx] + d2E_x/dy2 + k0^2 eps E_x = beta^2 E_x
    TM (E_y dominant):  d2E_y/dx2 + d/dy[(1/eps) d(eps E_y)/dy] + k0^2 eps E_y = beta^2 E_y

(eps at the half nodes is the mean of its two neighbours), shift-inverted at
k0^2 max(eps) so the first modes returned are the most strongly guided.  The
remaining components follow from div(eps E) = 0 and Faraday's law for fields
~ exp(i(w t - beta z)), z the propagation axis, e.g. for TE
E_z = -i/(beta eps) d(eps E_x)/dx, H_y = beta E_x/(w mu0) - i/(w mu0) dE_z/dx.
Profiles are scaled to max |dominant E| = 1.  The solver is host set-up
(scipy), run once per source; the time step never calls it.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np
from scipy import sparse
from scipy.sparse import linalg as spla

C0 = 299792458.0
MU0 = 4 * np.pi * 1e-7


@dataclass
class WaveguideMode:
    mode_number: int
    neff: complex
    frequency: float
    wavelength: float
    Ex: np.ndarray
    Ey: np.ndarray
    Ez: np.ndarray
    Hx: np.ndarray
    Hy: np.ndarray
    Hz: np.ndarray
    x: np.ndarray
    y: np.ndarray
    power: float = 1.0


def _second_difference(n: int, h: float) -> sparse.csr_matrix:
    """d2/ds2 on n nodes with zero field beyond both ends."""
    main = np.full(n, -2.0 / h**2)
    off = np.full(n - 1, 1.0 / h**2)
    return sparse.diags([off, main, off], [-1, 0, 1], format="csr")


def _weighted_difference(eps_line: np.ndarray, h: float) -> sparse.csr_matrix:
    """v -> d/ds[(1/eps) d(eps v)/ds] on one line of nodes (zero beyond the
    ends, eps continued as the end value there)."""
    e = np.asarray(eps_line, dtype=np.float64)
    n = len(e)
    ep = np.append(e[1:], e[-1])           # eps at i+1
    em = np.insert(e[:-1], 0, e[0])        # eps at i-1
    e_hp = 0.5 * (e + ep)                  # eps at i+1/2
    e_hm = 0.5 * (e + em)                  # eps at i-1/2
    diag = -(e / e_hp + e / e_hm) / h**2
    up = (ep / e_hp)[:-1] / h**2           # coefficient of v[i+1] in row i
    lo = (em / e_hm)[1:] / h**2            # coefficient of v[i-1] in row i
    return sparse.diags([lo, diag, up], [-1, 0, 1], format="csr")


class ModeSolver:
    def __init__(self, wavelength: float, x: np.ndarray, y: np.ndarray, epsilon: np.ndarray):
        self.wavelength = wavelength
        self.frequency = C0 / wavelength
        self.k0 = 2 * np.pi / wavelength
        self.x = np.asarray(x, dtype=np.float64)
        self.y = np.asarray(y, dtype=np.float64)
        self.epsilon = np.asarray(epsilon, dtype=np.float64)
        self.dx = self.x[1] - self.x[0] if len(self.x) > 1 else 1.0
        self.dy = self.y[1] - self.y[0] if len(self.y) > 1 else 1.0
        self.nx = len(self.x)
        self.ny = len(self.y)
        if self.epsilon.shape != (self.nx, self.ny):
            raise ValueError(f"epsilon must be ({self.nx}, {self.ny}), got {self.epsilon.shape}")
        self.modes: list = []

    # ------------------------------------------------------------------ operator
    def operator(self, mode_type: str = "TE") -> sparse.csr_matrix:
        """The semi-vectorial operator on the flattened profile (x fastest:
        node (i, j) at i + nx j)."""
        nx, ny, e = self.nx, self.ny, self.epsilon
        Ix, Iy = sparse.identity(nx, format="csr"), sparse.identity(ny, format="csr")
        t = mode_type.upper()
        if t == "TE":
            # x part weighted per row j, y part plain
            blocks = [_weighted_difference(e[:, j], self.dx) for j in range(ny)]
            Axx = sparse.block_diag(blocks, format="csr")
            Ayy = sparse.kron(_second_difference(ny, self.dy), Ix, format="csr")
        elif t == "TM":
            Axx = sparse.kron(Iy, _second_difference(nx, self.dx), format="csr")
            # y part weighted per column i: permute to y-fastest and back
            blocks = [_weighted_difference(e[i, :], self.dy) for i in range(nx)]
            B = sparse.block_diag(blocks, format="csr")          # y fastest
            idx = np.arange(nx * ny).reshape(ny, nx).T.ravel()   # y-fastest -> x-fastest
            P = sparse.csr_matrix((np.ones(nx * ny), (np.arange(nx * ny), idx)),
                                  shape=(nx * ny, nx * ny))
            Ayy = (P.T @ B @ P).tocsr()
        else:
            raise ValueError(f"Unknown mode type: {mode_type}")
        K = sparse.diags(self.k0**2 * e.T.ravel())              # x fastest
        return (Axx + Ayy + K).tocsr()

    # ------------------------------------------------------------------ solve
    def solve(self, num_modes: int = 1, mode_type: str = "TE",
              beta_guess: Optional[float] = None, guided_only: bool = True) -> list:
        """The ``num_modes`` most strongly guided modes, n_eff descending
        (``guided_only``: k0^2 min(eps) < beta^2 < k0^2 max(eps), as the
        reference filters; False keeps every propagating mode of the closed
        window, beta^2 > 0)."""
        t = mode_type.upper()
        if t == "VECTOR":
            raise NotImplementedError("full-vector modes are not built; use 'TE' or 'TM'")
        A = self.operator(t)
        n = self.nx * self.ny
        emax, emin = float(self.epsilon.max()), float(self.epsilon.min())
        sigma = (beta_guess**2 if beta_guess is not None else self.k0**2 * emax)
        k = int(min(max(num_modes + 2, 2 * num_modes), n - 2))
        vals, vecs = spla.eigs(A, k=k, sigma=sigma, which="LM")
        order = np.argsort(-vals.real)
        modes = []
        for m in order:
            b2 = vals[m].real
            lo = self.k0**2 * emin if guided_only else 0.0
            if not (lo < b2 < self.k0**2 * emax):
                continue
            beta = np.sqrt(b2)
            v = vecs[:, m]
            v = v / v[np.argmax(np.abs(v))]      # real eigenvector, largest entry 1
            f = v.real.reshape(self.ny, self.nx).T                    # (nx, ny)
            modes.append(self._fields(len(modes), beta, f, t))
            if len(modes) == num_modes:
                break
        self.modes = modes
        return modes

    def _fields(self, number, beta, f, t) -> WaveguideMode:
        w = 2 * np.pi * self.frequency
        e = self.epsilon
        gx = lambda a: np.gradient(a, self.dx, axis=0) if self.nx > 1 else np.zeros_like(a)
        gy = lambda a: np.gradient(a, self.dy, axis=1) if self.ny > 1 else np.zeros_like(a)
        zero = np.zeros_like(f, dtype=complex)
        if t == "TE":
            Ex = f.astype(complex)
            Ey = zero.copy()
            Ez = -1j / (beta * e) * gx(e * f)
            Hx = 1j / (w * MU0) * gy(Ez)
            Hy = beta / (w * MU0) * Ex - 1j / (w * MU0) * gx(Ez)
            Hz = -1j / (w * MU0) * gy(Ex)
        else:
            Ex = zero.copy()
            Ey = f.astype(complex)
            Ez = -1j / (beta * e) * gy(e * f)
            Hx = 1j / (w * MU0) * gy(Ez) - beta / (w * MU0) * Ey
            Hy = -1j / (w * MU0) * gx(Ez)
            Hz = 1j / (w * MU0) * gx(Ey)
        S = 0.5 * np.real(Ex * np.conj(Hy) - Ey * np.conj(Hx))
        power = float(S.sum() * self.dx * self.dy)
        return WaveguideMode(mode_number=number, neff=complex(beta / self.k0),
                             frequency=self.frequency, wavelength=self.wavelength,
                             Ex=Ex, Ey=Ey, Ez=Ez, Hx=Hx, Hy=Hy, Hz=Hz, x=self.x, y=self.y,
                             power=power)

    def get_mode(self, mode_number: int = 0) -> WaveguideMode:
        if mode_number >= len(self.modes):
            raise IndexError(
                f"Mode {mode_number} not found. Only {len(self.modes)} modes calculated.")
        return self.modes[mode_number]
