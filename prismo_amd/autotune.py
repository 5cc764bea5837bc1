"""Autotuners for the fused Yee step's launch shape and for the batched
2-D step (``tune_batch``).

Times each candidate (cells per lane, rows per workgroup, z planes per
workgroup, tile order, rows whole or in two segments, and one or two steps
per launch) for a few steps with
HIP events on the library's own stream, keeps the fastest per step, and
restores the simulation state it started from:
fields, CPML memory and dispersive state are snapshotted on the device first,
so tuning leaves a simulation exactly where it was.  Results are cached per
(grid, precision, coefficient storage, CPML, ADE) in memory and, when
``PRISMO_AMD_TUNE_CACHE`` names a file, in that JSON file.
"""

from __future__ import annotations

import json
import os
from itertools import product

_CACHE: dict = {}

CANDIDATES = {
    "vec": (1, 2),
    "ty": (2, 3, 4, 5, 7),
    "kchunk": (8, 12, 16, 24),
    "order": (2, 3),
}
# two steps per launch: rows per workgroup and cells per lane are fixed by
# the kernel (4 rows, the padded row width), z planes and tile order vary
CANDIDATES2 = {
    "kchunk": (16, 24, 32, 48, 64),
    "order": (2, 3),
}


# 2-D one-pass step of one simulation: rows per tile
CANDIDATES_2D = (1, 2, 3, 4, 8)

# batched 2-D step: rows per tile (the plan's chunk), waves per workgroup,
# XCD-grouped workgroup order
BATCH_CANDIDATES = {
    "chunk": (4, 8, 16, 32),
    "waves": (2, 4, 8),
    "xcd": (0, 1),
}


def _key(eng) -> str:
    p = eng.plan
    return json.dumps([p.nx, p.ny, p.nz, int(p.precision), int(p.coeff_kind),
                       list(p.pml), int(p.n_ade)])


def _load_file():
    path = os.environ.get("PRISMO_AMD_TUNE_CACHE")
    if path and os.path.exists(path):
        try:
            _CACHE.update(json.load(open(path)))
        except (OSError, ValueError):
            pass


def _save_file():
    path = os.environ.get("PRISMO_AMD_TUNE_CACHE")
    if path:
        try:
            json.dump(_CACHE, open(path, "w"))
        except OSError:
            pass


def _state_tensors(eng):
    """Every device tensor a step reads and writes in place."""
    out = [eng.fields[c] for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")]
    out += eng._psi_tensors()
    for P, Pm in getattr(eng, "ade_state", []):
        out.append(P)
        if Pm is not None:
            out.append(Pm)
    return out


def tune(eng, steps: int = 6, warmup: int = 2, candidates: dict | None = None,
         force: bool = False) -> dict:
    """Pick and apply the fastest fused-step launch shape for ``eng``.

    Returns {"vec", "ty", "kchunk", "order", "ms"} (``fused`` False when the
    plan cannot take the fused step; then nothing is changed)."""
    import torch

    if not getattr(eng, "fused", False):
        return {"fused": False}
    _load_file()
    key = _key(eng)
    if key in _CACHE and not force:
        best = _CACHE[key]
        eng.ctx.set_row_split(best.get("xs", 0))
        eng.ctx.set_tuning(best.get("steps", 1), best["ty"], best["kchunk"], best["vec"],
                           best["order"])
        return dict(best, fused=True, cached=True)
    cand = candidates or CANDIDATES
    splits = cand.get("xs", (1, 2))
    if getattr(eng, "is_2d", False):
        # the 2-D pass: rows per tile (kchunk) only
        shapes = [(1, 0, 2, kc, 2, 1) for kc in cand.get("rows_2d", CANDIDATES_2D)]
    else:
        shapes = [(1, vec, ty, kc, order, xs) for xs, vec, ty, kc, order in
                  product(splits, cand["vec"], cand["ty"], cand["kchunk"], cand["order"])]
    if not getattr(eng, "is_2d", False) and (candidates is None or "two_step" in candidates):
        c2 = (candidates or {}).get("two_step", CANDIDATES2)
        shapes += [(2, 0, 4, kc, order, 1) for kc, order in product(c2["kchunk"], c2["order"])]
    saved = [t.clone() for t in _state_tensors(eng)]
    best = None
    seen = set()
    try:
        for nstep, vec, ty, kc, order, xs in shapes:
            try:
                eng.ctx.set_row_split(xs)
                eng.ctx.set_tuning(nstep, ty, kc, vec, order)
                if eng.ctx.fused_steps() != nstep:
                    continue
                if nstep == 1:
                    # time each launch shape once, under the shape it runs
                    # (a request the kernel cannot take falls back)
                    shape = eng.ctx.fused_shape()
                    rows = kc if getattr(eng, "is_2d", False) else ty
                    if shape is None or (vec and shape[0] != vec) or shape[2] != rows \
                            or shape[3] != xs or (shape, kc, order) in seen:
                        continue
                    seen.add((shape, kc, order))
                # whole pairs of passes (4 steps) keep the two-step path on
                run = lambda n: eng.ctx.run(0, -(-n // 4) * 4, 0)
                run(warmup)
                eng.ctx.set_timing(True)
                run(steps)
                ms, n = eng.ctx.timing(3 if nstep == 1 else 5)
                eng.ctx.set_timing(False)
            except (ValueError, NotImplementedError):
                eng.ctx.set_timing(False)
                continue
            if n and (best is None or ms / nstep < best["ms"]):
                best = {"steps": nstep, "vec": vec, "ty": ty, "kchunk": kc, "order": order,
                        "xs": xs, "ms": ms / nstep, "ms_launch": ms}
    finally:
        for t, s in zip(_state_tensors(eng), saved):
            t.copy_(s)
        torch.cuda.synchronize(eng.device)
    if best is None:
        eng.ctx.set_row_split(0)
        eng.ctx.set_tuning(True, 4, 8, 0, 2)
        return {"fused": False}
    eng.ctx.set_row_split(best["xs"])
    eng.ctx.set_tuning(best["steps"], best["ty"], best["kchunk"], best["vec"], best["order"])
    _CACHE[key] = best
    _save_file()
    return dict(best, fused=True, cached=False)


def _batch_key(eng) -> str:
    # designs per batch rounded up to a power of two: a sweep's groups of
    # one size share a pick whatever the design sizes in them
    n = 1 << max(0, (len(eng.sims) - 1).bit_length())
    return json.dumps(["batch", int(eng.precision), int(eng.coeff_kind), n])


def tune_batch(eng, steps: int = 16, candidates: dict | None = None,
               force: bool = False) -> dict:
    """Pick and apply the fastest launch shape of the batched step for the
    configured, not yet run ``BatchEngine`` ``eng``: each candidate (rows per
    tile, waves per workgroup, workgroup order) runs ``steps`` eager steps
    timed with HIP events on the engine's stream; then the engine's state is
    cleared back to the zeros a run starts from and it is re-planned with
    the pick.  Results are bitwise the same under every shape (each cell's
    arithmetic does not depend on the tiling); cached per (precision,
    coefficient storage, designs per batch) like ``tune``.

    Returns {"chunk", "waves", "xcd", "ms", "cached"}."""
    _load_file()
    key = _batch_key(eng)
    cand = candidates or BATCH_CANDIDATES
    if key in _CACHE and not force:
        best = dict(_CACHE[key], cached=True)
    else:
        n = max(1, min(steps, eng.table_len))
        best = None
        try:
            for i, chunk in enumerate(cand["chunk"]):
                for waves, xcd in product(cand["waves"], cand["xcd"]):
                    eng.reconfigure(chunk)
                    eng.ctx.set_tuning(waves, xcd)
                    if i == 0 and best is None:
                        eng.ctx.run(min(2, n))    # kernel code loaded, untimed
                        eng.reconfigure(chunk)
                    eng.ctx.set_timing(True)
                    eng.ctx.run(n)
                    ms, cnt = eng.ctx.timing(0)
                    eng.ctx.set_timing(False)
                    if cnt and (best is None or ms < best["ms"]):
                        best = {"chunk": chunk, "waves": waves, "xcd": xcd, "ms": ms}
        finally:
            eng.ctx.set_timing(False)
            eng.ctx.sync()
            eng._zero_state()
        _CACHE[key] = best
        _save_file()
        best = dict(best, cached=False)
    eng.reconfigure(best["chunk"])
    eng.ctx.set_tuning(best["waves"], best["xcd"])
    return best
