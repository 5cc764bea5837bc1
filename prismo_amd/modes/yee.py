"""Guided modes of the time-stepping grid itself, for mode injection.

A mode launched into the FDTD grid only leaves a total-field /
scattered-field plane in one direction if it is an eigenmode of the grid's
own discrete operators.  This solver writes the 3-D stencil of the step
library on one cross-section, with (u, v, w) = the two transverse axes in
cyclic order and the propagation axis: E from forward differences of H, H
from backward differences of E (index space, zero beyond the plane), each
component with its own permittivity sample, fields ~ exp(i(w t - beta w)).
Along w the staggered difference of exp(-i beta w) is exactly -i beta~ with
beta~ = (2/dw) sin(beta dw / 2), and in time the leapfrog gives
w~ = (2/dt) sin(w dt / 2); with those, eliminating E_w and H_w leaves

    beta~^2 e = P e,  e = (E_u, E_v) on the plane,

    A_v e = beta~ H_v = w~ eps_u E_u - (1/(w~ mu0)) Dv+ (Du- E_v - Dv- E_u)
    A_u e = beta~ H_u = -w~ eps_v E_v - (1/(w~ mu0)) Du+ (Du- E_v - Dv- E_u)
    Q = Du+ A_v - Dv+ A_u
    P e = [ w~ mu0 A_v + (1/w~) Du- eps_w^-1 Q ;  -w~ mu0 A_u + (1/w~) Dv- eps_w^-1 Q ]

which reduces to the Helmholtz operator k0^2 eps + laplacian for a uniform
medium.  The eigenvector's (E_u, E_v) and H = A e / beta~ are the grid
samples a TF/SF sheet needs, at the indices the stencil pairs them with.
"""

from __future__ import annotations

import numpy as np
from scipy import sparse
from scipy.sparse import linalg as spla

C0 = 299792458.0
MU0 = 4 * np.pi * 1e-7
EPS0 = 8.854187817e-12


def _dplus(n_u, n_v, h, along_u):
    """Forward difference (f[i+1] - f[i]) / h on the (n_u, n_v) index grid,
    u fastest, zero beyond the grid."""
    n = n_u if along_u else n_v
    d1 = sparse.diags([-np.ones(n), np.ones(n - 1)], [0, 1], format="csr") / h
    if along_u:
        return sparse.kron(sparse.identity(n_v), d1, format="csr")
    return sparse.kron(d1, sparse.identity(n_u), format="csr")


def yee_modes(eps_u, eps_v, eps_w, du, dv, dw, omega, dt, n_modes=4, n_guess=None):
    """Discrete guided modes of one cross-section.

    eps_u, eps_v, eps_w: relative permittivity at the E_u, E_v, E_w samples,
    each (n_u, n_v).  Returns [(n_eff, Eu, Ev, Hu, Hv)], n_eff descending,
    fields (n_u, n_v) real, scaled to max |E_t| = 1; n_eff from
    beta = (2/dw) asin(beta~ dw / 2)."""
    eu, ev, ew = (np.asarray(a, dtype=np.float64) for a in (eps_u, eps_v, eps_w))
    n_u, n_v = eu.shape
    m = n_u * n_v
    wt = (2.0 / dt) * np.sin(omega * dt / 2)
    Dup, Dvp = _dplus(n_u, n_v, du, True), _dplus(n_u, n_v, dv, False)
    Dum, Dvm = -Dup.T.tocsr(), -Dvp.T.tocsr()
    flat = lambda a: a.T.ravel()          # (n_u, n_v) -> u fastest
    Eu_ = sparse.hstack([sparse.identity(m), sparse.csr_matrix((m, m))], format="csr")
    Ev_ = sparse.hstack([sparse.csr_matrix((m, m)), sparse.identity(m)], format="csr")
    curl_w = (Dum @ Ev_ - Dvm @ Eu_)      # Du- E_v - Dv- E_u
    c1 = 1.0 / (wt * MU0)
    Av = wt * EPS0 * sparse.diags(flat(eu)) @ Eu_ - c1 * (Dvp @ curl_w)
    Au = -wt * EPS0 * sparse.diags(flat(ev)) @ Ev_ - c1 * (Dup @ curl_w)
    Q = Dup @ Av - Dvp @ Au
    iw = sparse.diags(1.0 / (EPS0 * flat(ew)))
    P = sparse.vstack([wt * MU0 * Av + (1.0 / wt) * (Dum @ iw @ Q),
                       -wt * MU0 * Au + (1.0 / wt) * (Dvm @ iw @ Q)], format="csc")
    k0t = wt / C0
    emax = max(eu.max(), ev.max())
    sigma = (k0t * (n_guess if n_guess else np.sqrt(emax)))**2
    k = int(min(n_modes + 2, 2 * m - 2))
    vals, vecs = spla.eigs(P, k=k, sigma=sigma, which="LM")
    out = []
    for idx in np.argsort(-vals.real):
        b2 = vals[idx].real
        if b2 <= 0 or abs(vals[idx].imag) > 1e-6 * abs(b2):
            continue
        bt = np.sqrt(b2)
        x = bt * dw / 2
        if x >= 1.0:
            continue
        beta = (2.0 / dw) * np.arcsin(x)
        v = vecs[:, idx]
        v = v / v[np.argmax(np.abs(v))]
        v = v.real
        hu, hv = (Au @ v) / bt, (Av @ v) / bt
        shape = lambda a: a.reshape(n_v, n_u).T
        Eu, Ev, Hu, Hv = shape(v[:m]), shape(v[m:]), shape(hu), shape(hv)
        s = np.sqrt(np.max(Eu**2 + Ev**2))
        out.append((beta / (omega / C0), Eu / s, Ev / s, Hu / s, Hv / s))
    return out
