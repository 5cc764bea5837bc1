"""GPU: port |S21| / |S31| of batched directional couplers against values
the physics fixes (the flux post-processing of analysis/sparameters.py and
workloads.coupler_analyze on fields from the batched kernel).

* Straight guide (the normalisation run): the forward flux through the
  output window is at most the flux through the input window and at least
  97 % of it (a lossless guide; what is lost is radiation the line source
  launches outside the guided mode, which leaves through the y CPML between
  the planes); almost nothing crosses the empty cross window.
* Coupler, normalised to its straight guide: |S21|^2 + |S31|^2 within
  [0.98, 1.01] (lossless; the margin above 1 is source radiation that the
  second strip catches instead of the CPML).
* The cross fraction |S31|^2 / (|S21|^2 + |S31|^2) against coupled-mode
  theory: the input strip's mode, at the start of the second strip, is
  expanded in the two supermodes of the coupled section, each solved on the
  grid's own discrete operator (y second difference, the stencil's
  dispersion in x and in time) by a dense eigen-solve, propagated over the
  coupled length and split at the mid-line by Poynting flux.  Agreement
  within 0.07 absolute at all 64 frequencies, on designs coupling 35-95 %
  across the band.  The model leaves out the abrupt start of the second
  strip and the radiation; the grid is its own (the same averaged
  permittivity at the Ez samples), and one cell (20 nm) more or less gap
  moves the fraction by 0.15-0.2 on these designs, so the bar separates a
  correct grid from one misplaced cell.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
C0 = 299792458.0
D = 20e-9
STEPS = 6000
DESIGNS = [{"gap": 100e-9, "length": 6e-6, "width": 450e-9, "clad_eps": 2.085},
           {"gap": 100e-9, "length": 9e-6, "width": 450e-9, "clad_eps": 2.085},
           {"gap": 300e-9, "length": 3e-6, "width": 450e-9, "clad_eps": 2.085}]


def _ez_eps(eps_col):
    """Permittivity at the Ez samples between nodes j and j+1 (the harmonic
    mean: Cb is proportional to 1/eps and averaged)."""
    return 2.0 / (1.0 / eps_col[:-1] + 1.0 / eps_col[1:])


def _cmt_cross_fraction(p, sim, freqs):
    from scipy.linalg import eigh

    from prismo_amd import workloads as W

    eps = np.asarray(sim._material_arrays["eps_rel"])[:, :, 0]
    t = W.COUPLER_PML
    nxp, nyp = W.coupler_cells(p)
    x_out = nxp * D - 0.6e-6                 # output plane
    l_cpl = x_out - 1.5e-6                   # the second strip starts at 1.5 um
    i_in = t + int(round(1.0e-6 / D))        # single-strip column
    i_cp = t + int(round((1.5e-6 + x_out) / 2 / D))  # coupled column
    e1, e2 = _ez_eps(eps[i_in]), _ez_eps(eps[i_cp])
    n = e1.size
    y = (np.arange(n) + 0.5 - t) * D
    upper = y > nyp * D / 2
    lap = (np.diag(np.full(n, -2.0)) + np.diag(np.ones(n - 1), 1)
           + np.diag(np.ones(n - 1), -1)) / D**2
    dt = sim.dt
    out = []
    for f in freqs:
        k2 = (2.0 / (C0 * dt) * np.sin(np.pi * f * dt)) ** 2

        def modes(e, m):
            lam, v = eigh(lap + np.diag(k2 * e))
            idx = np.argsort(lam)[::-1][:m]
            return np.sqrt(lam[idx]), v[:, idx]   # (2/D) sin(beta D / 2), profiles

        kt1, v1 = modes(e1, 1)
        kt2, v2 = modes(e2, 2)
        beta = 2.0 / D * np.arcsin(kt2 * D / 2)
        a = v2.T @ v1[:, 0]
        ez = v2 @ (a * np.exp(-1j * beta * l_cpl))
        hy = v2 @ (a * kt2 * np.exp(-1j * beta * l_cpl))
        sx = np.real(ez * np.conj(hy))
        out.append(sx[upper].sum() / sx.sum())
    return np.array(out)


def test_coupler_s21_power_and_coupled_mode_theory(mi355x):
    from prismo_amd import workloads as W
    from prismo_amd.core.batch import BatchEngine

    sims = []
    for p in DESIGNS:
        sims.append(W.coupler_sim(p, dtype=np.float64, backend=W.host_backend()))
        sims.append(W.coupler_sim(p, dtype=np.float64, backend=W.host_backend(), straight=True))
    eng = BatchEngine(sims, dtype=np.float64, chunk=16, sample_T=128)
    eng.run_steps(STEPS)
    assert not any(eng.diverged)
    freqs = np.asarray(W.COUPLER_FREQS)
    checks = []
    for k, p in enumerate(DESIGNS):
        dev, st = sims[2 * k], sims[2 * k + 1]
        nrm = W.coupler_norm(st)
        thru_in = nrm["p_thru"] / nrm["p_in"]
        cross_in = nrm["p_cross"] / nrm["p_in"]
        r = W.coupler_analyze(dev, p, norm=nrm)
        total = r["s21"] ** 2 + r["s31"] ** 2
        frac = r["s31"] ** 2 / total
        cmt = _cmt_cross_fraction(p, dev, freqs)
        print(f"gap {p['gap'] * 1e9:.0f} nm L {p['length'] * 1e6:.0f} um: straight thru/in "
              f"[{thru_in.min():.4f}, {thru_in.max():.4f}] cross/in max {cross_in.max():.4f}; "
              f"|S21|^2+|S31|^2 [{total.min():.4f}, {total.max():.4f}]; cross fraction "
              f"{frac[0]:.3f}..{frac[-1]:.3f} vs CMT {cmt[0]:.3f}..{cmt[-1]:.3f}, "
              f"max diff {np.abs(frac - cmt).max():.4f}", flush=True)
        r_in = W.coupler_analyze(dev, p)
        checks.append((p, thru_in, cross_in, total, frac, cmt, r, r_in, nrm))
    for p, thru_in, cross_in, total, frac, cmt, r, r_in, nrm in checks:
        assert np.all(thru_in <= 1.0) and np.all(thru_in >= 0.97)
        assert np.all(cross_in <= 0.01)
        assert np.all(total >= 0.98) and np.all(total <= 1.01)
        assert np.abs(frac - cmt).max() <= 0.07
        # the input-port normalisation (the reference's incident power) and
        # the straight-guide one differ by the straight guide's thru/in ratio
        assert np.allclose(r_in["s21"] ** 2 * nrm["p_in"], r["s21"] ** 2 * nrm["p_out"])
        if p["gap"] == 100e-9:
            # strongly coupled: the comparison spans a large transfer
            assert frac.max() >= 0.4
