"""Surrogate-screened design search over directional couplers.

1. Spectra for ``n_train`` seeded designs (gap, length, strip width,
   cladding eps) from the batched FDTD (prismo_amd.core.batch).
2. An ensemble of ``SpectralMLP`` maps the 4 normalised parameters to the
   64-frequency complex through and cross transmission (re, im: 256
   outputs), trained by the device kernels.
3. ``n_screen`` seeded candidates are scored by the ensemble; the top ``k``
   by the design objective and every candidate whose predicted objective
   is uncertain (ensemble spread above ``tau`` near the acceptance band)
   are re-run by batched FDTD, which decides.  Only designs the FDTD runs accept
   count.

Objective (a 3 dB coupler at 193 THz): |S21|^2 and |S31|^2 both within
``band`` of 0.5 and |S21|^2 + |S31|^2 >= ``min_total``.
"""

from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from .. import workloads as W
from .mlp import SurrogateEnsemble

RANGES = {"gap": (100e-9, 400e-9), "length": (2e-6, 12e-6), "width": (400e-9, 500e-9),
          "clad_eps": (2.0, 2.2)}
KEYS = ("gap", "length", "width", "clad_eps")


def encode(designs) -> np.ndarray:
    return np.array([[2 * (d[k] - RANGES[k][0]) / (RANGES[k][1] - RANGES[k][0]) - 1 for k in KEYS]
                     for d in designs], dtype=np.float32)


def targets(results) -> np.ndarray:
    """(n, 256): re/im of the through and cross transmission, 64 frequencies
    each, with the through port's phase taken as the reference (the absolute
    phase only moves with the arbitrary reference plane): through =
    (|S21|, 0), cross = |S31| exp(i (phase_cross - phase_through))."""
    rows = []
    for r in results:
        ref = np.exp(-1j * np.angle(r["t21"]))
        a = r["t21"] * ref
        b = r["t31"] * ref
        rows.append(np.concatenate([a.real, a.imag, b.real, b.imag]))
    return np.array(rows, dtype=np.float32)


def magnitudes(Y) -> tuple:
    Y = np.asarray(Y, dtype=np.float64)
    nf = Y.shape[1] // 4
    s21 = np.hypot(Y[:, 0:nf], Y[:, nf:2 * nf])
    s31 = np.hypot(Y[:, 2 * nf:3 * nf], Y[:, 3 * nf:4 * nf])
    return s21, s31


@dataclass
class Objective:
    f_index: int = int(np.argmin(np.abs(W.COUPLER_FREQS - 193e12)))
    band: float = 0.1
    min_total: float = 0.8

    def score(self, s21, s31):
        """Lower is better: distance of the split from 50/50 plus the
        power lost below ``min_total``."""
        a = s21[:, self.f_index] ** 2
        b = s31[:, self.f_index] ** 2
        return np.abs(a - 0.5) + np.abs(b - 0.5) + np.maximum(self.min_total - (a + b), 0.0)

    def accept(self, s21, s31):
        a = s21[:, self.f_index] ** 2
        b = s31[:, self.f_index] ** 2
        return (np.abs(a - 0.5) <= self.band) & (np.abs(b - 0.5) <= self.band) & (
            a + b >= self.min_total)


def fdtd(designs, steps=W.COUPLER_STEPS, max_batch=64, dtype=np.float32, storage="fp16"):
    """Batched FDTD check: coupler_analyze results per design, through the
    pipelined batched sweep (groups of ``max_batch``, the first of 16: one
    group is built and analysed on the host while the previous one steps on
    the GPU)."""
    from ..optimization.sweep import BatchedRunner, ParameterSweep

    runner = BatchedRunner(build=lambda p: W.coupler_sim(p, backend=W.host_backend()),
                           analyze=lambda sim, p: W.coupler_analyze(sim), steps=steps,
                           dtype=dtype, coefficient_storage=storage, max_batch=max_batch,
                           engine_kw={"chunk": "auto"}, first_batch=min(16, max_batch))
    res = ParameterSweep.from_samples(designs, batched=runner).run()
    bad = [r for r in res if r["status"] != "success"]
    if bad:
        raise RuntimeError(f"FDTD check: {len(bad)} designs failed, first: {bad[0]['error']}")
    return res


def run(n_train=4096, n_screen=65536, top_k=64, tau=0.05, members=4, epochs=150,
        steps=W.COUPLER_STEPS, seed=W.SEED, log=print) -> dict:
    import torch

    obj = Objective()
    timings = {}
    t0 = time.perf_counter()
    train = W.coupler_designs(n_train, seed=seed, width=True)
    res = fdtd(train, steps=steps)
    torch.cuda.synchronize()
    timings["train_fdtd_s"] = time.perf_counter() - t0
    log(f"[surrogate] FDTD spectra for {n_train} designs in {timings['train_fdtd_s']:.1f} s")

    s21t = np.array([r["s21"] for r in res])
    s31t = np.array([r["s31"] for r in res])
    fi0 = obj.f_index
    log(f"[surrogate] training set at 193 THz: |S21|^2 mean {np.mean(s21t[:, fi0] ** 2):.3f}, "
        f"|S31|^2 mean {np.mean(s31t[:, fi0] ** 2):.3f}, total mean "
        f"{np.mean(s21t[:, fi0] ** 2 + s31t[:, fi0] ** 2):.3f}")
    X = torch.from_numpy(encode(train)).cuda()
    T = torch.from_numpy(targets(res)).cuda()
    t1 = time.perf_counter()
    ens = SurrogateEnsemble(n_members=members, n_in=4, n_out=T.shape[1], hidden=256, depth=4)
    hist = ens.fit(X, T, epochs=epochs, batch=256, lr=2e-3)
    torch.cuda.synchronize()
    timings["train_s"] = time.perf_counter() - t1
    log(f"[surrogate] trained {members} members x {epochs} epochs in {timings['train_s']:.1f} s, "
        f"final loss {[round(h[-1], 5) for h in hist]}")

    t2 = time.perf_counter()
    cand = W.coupler_designs(n_screen, seed=seed + 1, width=True)
    Xc = torch.from_numpy(encode(cand)).cuda()
    mean, std = ens.predict(Xc)
    torch.cuda.synchronize()
    timings["screen_s"] = time.perf_counter() - t2
    mu = mean.cpu().numpy()
    sd = std.cpu().numpy()
    s21p, s31p = magnitudes(mu)
    score = obj.score(s21p, s31p)
    # spread of the objective's inputs at the design frequency
    nf = mu.shape[1] // 4
    fi = obj.f_index
    unc = np.max(sd[:, [fi, nf + fi, 2 * nf + fi, 3 * nf + fi]], axis=1)
    order = np.argsort(score)
    pick = list(order[:top_k])
    near = np.nonzero((score < 3 * obj.band) & (unc > tau))[0]
    pick += [i for i in near[np.argsort(score[near])] if i not in set(pick)][:top_k]
    log(f"[surrogate] screened {n_screen} in {timings['screen_s']:.2f} s; checking {len(pick)} by FDTD "
        f"({top_k} best + {len(pick) - top_k} uncertain)")

    t3 = time.perf_counter()
    checked = fdtd([cand[i] for i in pick], steps=steps)
    torch.cuda.synchronize()
    timings["fdtd_check_s"] = time.perf_counter() - t3
    s21f = np.array([r["s21"] for r in checked])
    s31f = np.array([r["s31"] for r in checked])
    ok = obj.accept(s21f, s31f)
    err21 = np.abs(s21p[pick] - s21f)
    err31 = np.abs(s31p[pick] - s31f)
    total = time.perf_counter() - t0

    # the method without the surrogate: FDTD on the candidates in their
    # seeded order, for the same FDTD budget
    rand_ok = obj.accept(np.array([r["s21"] for r in res]), np.array([r["s31"] for r in res]))
    return {
        "n_train": n_train, "n_screen": n_screen, "fdtd_checked": len(pick),
        "accepted": int(ok.sum()),
        "accept_rate_surrogate_picks": float(ok.mean()),
        "accept_rate_random_designs": float(rand_ok.mean()),
        "s21_abs_err_mean": float(err21.mean()), "s21_abs_err_max": float(err21.max()),
        "s31_abs_err_mean": float(err31.mean()),
        "s21_abs_err_at_193THz_mean": float(err21[:, obj.f_index].mean()),
        "timings_s": {k: round(v, 3) for k, v in timings.items()},
        "total_s": round(total, 3),
        "accepted_per_s": float(ok.sum() / total),
        "screened_per_s": float(n_screen / timings["screen_s"]),
        "final_loss": [float(h[-1]) for h in hist],
    }
