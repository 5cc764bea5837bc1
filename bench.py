"""Benchmark: 3-D Yee + CPML time step on one MI355X.

Workload (``voxel_3d``): 256^3 cells at 20 nm, 12-cell CPML on all six faces
(inside the 256^3), eps per 8^3 voxel block from {1.0, 2.085, 12.1}
(seed 20261202), Ez dipole at the centre with a 193 THz Gaussian pulse,
FP32 fields with FP16 Cb coefficients.  One step = H update + E update
(with CPML terms in the slabs) + source injection.

Prints one JSON line: Gcell-updates/s over the timed steps, the roofline of
the fused step kernel and a CPU baseline (the NumPy FP64 O1 oracle on the
host, a few steps of the same workload, one core).

Algorithmic bytes of one step (no reuse counted): 54 B per cell (6 FP32
components read and written, 3 FP16 Cb coefficients) plus 32 B per
(cell, CPML axis) inside the slabs (two E-side and two H-side FP32 psi
values read and written).  ``achieved`` divides that by the average duration
of the fused kernel, timed with HIP events on the library's own stream over
the timed steps.  ``traffic`` is the HBM bytes one launch moves as measured
by the rocprofv3 FETCH_SIZE / WRITE_SIZE counters (scaled by a copy-kernel
calibration), read from the committed profile summary for this workload
when it is present.

The same line carries ``sweep``: 256 seeded directional couplers (2-D TM,
200x120 .. 800x400 cells + 16-cell CPML, 4000 steps, 64-frequency ports)
advanced together in FP32 with FP16 Cb by the batched kernel, timed end to
end (build, run, |S21| analysis) in pipelined groups (``designs_per_s``)
and as one batch of all 256 (``designs_per_s_one_batch``), and on the GPU,
beside sequential plain FP64 runs of an evenly spaced sample of 32 of the
same designs through Simulation.run; ``max_abs_s21_err`` is the largest
|S21| difference between the two over that sample and all 64 frequencies.

``strip_2d``: the 2-D TM strip (540 x 290 cells, FP64, 3000 steps) in the
reference-arithmetic and textbook schemes on the GPU, beside the O0 NumPy
restatement of the reference's step on one host core (a bounded number of
steps); small grids are launch-bound, so it reports the microseconds per
step.  ``waveguide_3d``: the dispersive waveguide (480 x 240 x 160 cells,
Lorentz core, Drude pad, mode source, 64-frequency DFT plane, 6000 steps)
with FP32 fields and FP32 Cb, FP32 fields and FP16 Cb, and FP64: the
relative L2 difference of each FP32 run's plane spectra from FP64's, and
each step kernel's roofline (``waveguide_alg_bytes`` over its HIP-event
time).
``voxel_fp64``: the voxel workload with FP64 fields and coefficients (the
plain method), the same steps: throughput, its roofline (120 B per cell,
64 B per CPML cell-axis) and the relative L2 difference over all six
components from the FP32/FP16 run at the end (FP16 Cb quantisation
included).  ``--no-strip`` / ``--no-waveguide`` / ``--no-sweep`` /
``--no-fp64`` skip those parts.

``surrogate``: the surrogate-screened coupler search (4096 FDTD training
spectra, a 4-member BF16 MLP ensemble, 65536 screened designs, FDTD check
of the top 64 and the uncertain ones): FDTD-confirmed designs per second end
to end, the |S21| / |S31| error of the surrogate against the FDTD and the
time of each stage; ``--no-surrogate`` skips it, ``--surrogate`` runs only
it and prints its own JSON line.

``voxel_accuracy``: the FP32/FP16 voxel fields after 1000 steps against an
FP64 run with the same (FP16-rounded) Cb and against FP64 with exact Cb.

``--waveguide-only [--waveguide-steps N]`` and ``--fp64-only`` run only
those legs (short runs for counter collection).  ``--copy-probe`` runs only a
256 MiB device copy (the counter calibration
kernel); ``--batch-probe`` only 8 eager steps of the 256-coupler batch (for
counter runs on the batch kernel); ``--launch vec,ty,kchunk,order[,steps[,xs]]``
fixes the fused launch shape (one or two steps per launch, rows whole or in
two segments) instead of autotuning.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
ALG_BYTES_PER_CELL = 54.0  # 12 * 4 B fields + 3 * 2 B coefficients
ALG_BYTES_PER_PML_CELL_AXIS = 32.0  # 4 psi * 4 B, read and written
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_yee_fused_voxel256_traffic.md")


def algorithmic_bytes(plan) -> float:
    n = (plan.nx, plan.ny, plan.nz)
    cells = n[0] * n[1] * n[2]
    slab = 0
    for a in range(3):
        slab += (plan.pml[2 * a] + plan.pml[2 * a + 1]) * cells // n[a]
    return ALG_BYTES_PER_CELL * cells + ALG_BYTES_PER_PML_CELL_AXIS * slab


def voxel_fp64(n, steps, warmup, ref):
    """The same voxel workload with FP64 fields and FP64 coefficients (the
    plain method), the same number of steps: throughput, the fused kernel's
    roofline (120 B per cell + 64 B per CPML cell-axis), and the relative
    L2 distance over all six components between the FP32/FP16 fields
    ``ref`` and these, at the end of the run (it includes the FP16
    quantisation of Cb, not only FP32 arithmetic)."""
    import numpy as np
    import torch

    from prismo_amd.workloads import voxel_3d

    sim = voxel_3d(dtype=np.float64, coefficient_storage="auto", n=n)
    eng = sim.solver.updater.engine_for(sim.fields)
    tuned = sim.autotune()
    sim.run_steps(warmup)
    sim.prepare(steps)
    torch.cuda.synchronize()
    eng.ctx.set_timing(True, stride=max(1, min(10, steps // 20)))
    t0 = time.perf_counter()
    sim.run_steps(steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    f_ms, nf = eng.ctx.timing(3)
    eng.ctx.set_timing(False)
    p = eng.plan
    cells = p.nx * p.ny * p.nz
    slab = sum((p.pml[2 * a] + p.pml[2 * a + 1]) * cells // (p.nx, p.ny, p.nz)[a]
               for a in range(3))
    alg = 120 * cells + 64 * slab
    num = den = 0.0
    for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz") if ref is not None else ():
        f = sim.fields[c].double()
        num += float(((f - ref[c].double()) ** 2).sum())
        den += float((f ** 2).sum())
    out = {"gcells_per_s": round(cells * steps / el / 1e9, 3),
           "ms_per_step": round(el / steps * 1e3, 5),
           "fused": bool(nf),
           "launch": {k: tuned.get(k) for k in ("vec", "ty", "kchunk", "order", "xs")},
           "rel_l2_fp32_fp16_vs_fp64_6comp": (num / den) ** 0.5 if den > 0 else None}
    if nf:
        out["roofline"] = {"bound": "hbm", "kernel": "k_yee_fused<double>",
                           "achieved": round(alg / (f_ms * 1e-3) / 1e9, 1),
                           "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(alg / (f_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                           "alg_bytes_per_launch": alg, "kernel_ms": round(f_ms, 5)}
    del sim, eng
    torch.cuda.empty_cache()
    print(f"[bench] FP64 voxel: {out}", file=sys.stderr, flush=True)
    return out


ACC_STEPS = 1000


def voxel_accuracy(n, tuned, steps=ACC_STEPS):
    """The precision form against the plain method on the voxel workload
    over a fixed 1000 steps, whatever the timed run's length: FP32 fields
    with FP16 Cb (the timed launch shape) against FP64 fields fed the same
    FP16-rounded Cb (arithmetic error only), and against FP64 with the exact
    Cb (the quantisation included); relative L2 over all six components."""
    import numpy as np
    import torch

    from prismo_amd.workloads import voxel_3d

    comps = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
    s32 = voxel_3d(dtype=np.float32, coefficient_storage="fp16", n=n)
    e32 = s32.solver.updater.engine_for(s32.fields)
    if tuned.get("fused"):
        e32.ctx.set_tuning(tuned.get("steps", 1), tuned["ty"], tuned["kchunk"], tuned["vec"],
                           tuned["order"])
    s32.run_steps(steps)
    f32 = {c: s32.fields[c].double() for c in comps}
    cb_q = dict(e32.cb_host)     # the FP16-rounded Cb, in FP64
    del s32, e32
    torch.cuda.empty_cache()
    out = {"steps": steps}
    for tag, quant in (("vs_fp64_same_cb", True), ("vs_fp64_exact_cb", False)):
        s64 = voxel_3d(dtype=np.float64, coefficient_storage="auto", n=n)
        e64 = s64.solver.updater.engine_for(s64.fields)
        if quant:
            for c, v in cb_q.items():
                e64._padded(v, c, torch.float64, buf=e64.cb_dev[c])
        s64.run_steps(steps)
        num = den = 0.0
        for c in comps:
            f = s64.fields[c]
            num += float(((f32[c] - f) ** 2).sum())
            den += float((f ** 2).sum())
        out["rel_l2_" + tag] = (num / den) ** 0.5 if den > 0 else None
        del s64, e64
        torch.cuda.empty_cache()
    print(f"[bench] voxel accuracy over {steps} steps: {out}", file=sys.stderr, flush=True)
    return out


def measured_traffic(workload_key: str, launch: dict):
    """HBM bytes per fused launch from the committed counter summary, when
    it covers this workload and launch shape (else None)."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    want = tuple(launch.get(k) for k in ("vec", "ty", "kchunk", "order"))
    steps = launch.get("steps") or 1
    for line in open(TRAFFIC_FILE):
        if not line.startswith("traffic:"):
            continue
        vals = dict(kv.partition("=")[::2] for kv in line[len("traffic:"):].split())
        try:
            if vals.get("workload") != workload_key:
                continue
            if (tuple(int(x) for x in vals["launch"].split(",")) == want
                    and int(vals.get("steps", 1)) == steps
                    and int(vals.get("xs", 1)) == (launch.get("xs") or 1)):
                return float(vals["bytes"])
        except (KeyError, ValueError, TypeError):
            continue
    return None


def copy_probe():
    """256 MiB device-to-device copy, four times: the kernel whose counter
    readings calibrate FETCH_SIZE and WRITE_SIZE."""
    import torch

    a = torch.ones(256 << 18, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(4):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"[bench] copy probe: 4 x {a.numel() * 4} B copied", file=sys.stderr, flush=True)


def batch_probe(steps=8):
    """The 256 coupler designs of the sweep, one batch (8-row tiles), a few
    eager steps: one k_batch_tm launch per step for counter runs."""
    import numpy as np
    import torch

    from prismo_amd import workloads as W
    from prismo_amd.core.batch import BatchEngine

    sims = [W.coupler_sim(p, backend=W.host_backend()) for p in W.coupler_designs(256)]
    eng = BatchEngine(sims, dtype=np.float32, coefficient_storage="fp16", chunk=8, sample_T=128)
    eng.run_steps(steps, graph=False)
    torch.cuda.synchronize()
    print(f"[bench] batch probe: {steps} steps of 256 designs", file=sys.stderr, flush=True)


def cpu_baseline(n=256, steps=3):
    """O1 oracle (NumPy FP64, one core) on the same workload for a few steps."""
    import numpy as np

    from oracle import compat as O0
    from oracle.yee import YeeOracle
    from prismo_amd.workloads import voxel_eps

    pml = 12
    d = 20e-9
    eps = voxel_eps(n)
    c0 = 299792458.0
    dt = 0.99 / (c0 * np.sqrt(3 / d**2))
    Ca, Cb, Da, Db = O0.coefficients(eps, np.ones_like(eps), np.zeros_like(eps),
                                     np.zeros_like(eps), dt)
    cb = {"Ex": O0._ex(Cb), "Ey": O0._ey(Cb), "Ez": O0._ez(Cb)}
    del Ca, Da, Cb
    o1 = YeeOracle(n, n, n, (d, d, d), dt, cb, float(Db.flat[0]), pml=(pml,) * 6)
    o1.step()  # first step allocates the psi arrays; not timed
    t0 = time.perf_counter()
    for _ in range(steps):
        o1.step()
        o1.F["Ez"][n // 2 - 1, n // 2 - 1, n // 2] += 1e-3
    el = time.perf_counter() - t0
    return {
        "value": n**3 * steps / el / 1e9,
        "unit": "Gcell-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": f"NumPy FP64 textbook Yee+CPML oracle, {n}^3, {steps} steps "
                  f"after one untimed step ({el:.1f} s), host {os.cpu_count()} cpus",
    }


SWEEP_ALG_BYTES_PER_CELL = 26.0  # 2-D TM: Ez, Hx, Hy FP32 read + written, FP16 Cb
SWEEP_ALG_BYTES_PER_PML_CELL_AXIS = 16.0  # two FP32 psi read + written


def sweep_bench(n_designs=256, steps=4000, seq_sample=32, group=64, first=16):
    """Batched 256-coupler sweep through ParameterSweep's batched back end
    (FP32 fields, FP16 Cb, one launch per step over every design of a group
    of ``group``, the first group ``first`` designs so the GPU starts early,
    graph-replayed step loop, groups pipelined so the host builds and
    analyses while the GPU steps), timed end to end, against
    sequential plain FP64 runs of an evenly spaced sample of the same
    designs, after an untimed 64-step warm-up sweep of one group of each
    size.  Then all designs in one batch, the shape the workload is stated
    in (one launch per step over all 256, the step loop replayed as a HIP
    graph), timed end to end (build, run, |S21|) after an untimed 64-step
    run of that batch size (kernels loaded, launch shape tuned); the batch
    kernel's roofline is taken on that run (HIP events)."""
    import numpy as np
    import torch

    from prismo_amd import workloads as W
    from prismo_amd.core.batch import BatchEngine
    from prismo_amd.optimization.sweep import BatchedRunner, ParameterSweep

    designs = W.coupler_designs(n_designs)
    # rows per tile and launch shape from the autotuner (the untimed warm-up
    # below tunes each group size once; the timed run uses the cached picks)
    kw = {"chunk": "auto", "sample_T": 128}
    runner = BatchedRunner(build=lambda p: W.coupler_sim(p, backend=W.host_backend()),
                           analyze=lambda sim, p: W.coupler_analyze(sim), steps=steps,
                           dtype=np.float32, coefficient_storage="fp16", max_batch=group,
                           engine_kw=kw, first_batch=first)
    # untimed warm-up through the same path with the timed run's group
    # shapes (a group of `first`, then one of `group`), 64 steps: kernel code
    # loaded, graphs captured, device memory of those sizes in torch's pool
    warm = BatchedRunner(build=runner.build, analyze=runner.analyze, steps=64,
                         dtype=np.float32, coefficient_storage="fp16", max_batch=group,
                         engine_kw=kw, first_batch=first)
    ParameterSweep.from_samples(designs[:first + group], batched=warm).run()
    sweep = ParameterSweep.from_samples(designs, batched=runner)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = sweep.run()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    bad = [r for r in res if r["status"] != "success"]
    if bad:
        raise RuntimeError(f"sweep: {len(bad)} designs failed, first: {bad[0].get('error')}")
    rate = n_designs / (t1 - t0)
    print(f"[bench] sweep: {n_designs} designs end to end in {t1 - t0:.2f} s "
          f"(first group {first}, then groups of {group}, pipelined)", file=sys.stderr,
          flush=True)

    # all designs in one batch: one launch per step over all of them
    sims = [W.coupler_sim(p, backend=W.host_backend()) for p in designs]
    BatchEngine(sims, dtype=np.float32, coefficient_storage="fp16", **kw).run_steps(64)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    sims = [W.coupler_sim(p, backend=W.host_backend()) for p in designs]
    eng = BatchEngine(sims, dtype=np.float32, coefficient_storage="fp16", **kw)
    eng.run_steps(steps, graph=True, timing=True)
    one = [W.coupler_analyze(s) for s in sims]
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    one_rate = n_designs / (t3 - t2)
    one_err = max(float(np.max(np.abs(one[k]["s21"] - res[k]["s21"]))) for k in range(n_designs))
    print(f"[bench] sweep: {n_designs} designs in one batch end to end in {t3 - t2:.2f} s",
          file=sys.stderr, flush=True)
    tuning = eng.tuning
    g_ms, g_n = eng.timing(2)
    e_ms, e_n = eng.timing(0)
    f_ms, f_n = eng.timing(1)
    gpu_ms = g_ms * g_n + e_ms * e_n + f_ms * f_n
    cells = 0
    slab = 0
    for s in sims:
        nx, ny, _ = s.grid.dimensions
        cells += nx * ny
        t = W.COUPLER_PML
        slab += 2 * t * ny + 2 * t * nx
    del eng, sims
    alg = SWEEP_ALG_BYTES_PER_CELL * cells + SWEEP_ALG_BYTES_PER_PML_CELL_AXIS * slab
    step_ms = gpu_ms / steps

    idx = np.unique(np.linspace(0, n_designs - 1, seq_sample).round().astype(int))
    seq_t = 0.0
    err21 = err31 = 0.0
    for k in idx:
        torch.cuda.synchronize()
        a = time.perf_counter()
        sim = W.coupler_sim(designs[k], dtype=np.float64)
        sim.run_steps(steps)
        r = W.coupler_analyze(sim)
        torch.cuda.synchronize()
        seq_t += time.perf_counter() - a
        err21 = max(err21, float(np.max(np.abs(res[k]["s21"] - r["s21"]))))
        err31 = max(err31, float(np.max(np.abs(res[k]["s31"] - r["s31"]))))
        del sim
    seq_rate = len(idx) / seq_t
    print(f"[bench] sweep: FP64 sequential {len(idx)} designs in {seq_t:.2f} s", file=sys.stderr,
          flush=True)
    return {
        "designs": n_designs,
        "steps": steps,
        "designs_per_s": round(rate, 3),
        "group": group,
        "first_group": first,
        "designs_per_s_gpu_only": round(n_designs / (step_ms * steps / 1e3), 3),
        "designs_per_s_one_batch": round(one_rate, 3),
        "one_batch_vs_groups_max_abs_s21_diff": one_err,
        "fp64_sequential_designs_per_s": round(seq_rate, 4),
        "fp64_sample": int(len(idx)),
        "speedup": round(rate / seq_rate, 2),
        "max_abs_s21_err": err21,
        "max_abs_s31_err": err31,
        "cells_per_step": cells,
        "gcells_per_s": round(cells / (step_ms * 1e-3) / 1e9, 3),
        "roofline": {"bound": "hbm", "kernel": "k_batch_tm (one launch per step, all designs)",
                     "achieved": round(alg / (step_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(alg / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "alg_bytes_per_step": alg, "step_ms": round(step_ms, 5),
                     "steps_timed": steps, "launch": tuning},
        "precision": "fp32 fields, fp16 Cb, fp32 compensated spectra; baseline fp64",
    }


def _cpu_coupler(args):
    """One coupler design through the O0 NumPy restatement of the
    reference's step (its arithmetic, one core) with the design's plane-wave
    source and its three 64-frequency port monitors updated every step, as
    the reference's Simulation.run does; returns the seconds ``steps`` steps
    took."""
    k, steps = args
    import numpy as np

    from oracle import compat as O0
    from prismo_amd import workloads as W

    p = W.coupler_designs(256)[k]
    sim = W.coupler_sim(p, dtype=np.float64, backend=W.host_backend())
    nx, ny, nz = sim.grid.dimensions
    eps = np.asarray(sim.solver.updater.eps_rel).reshape(nx, ny, nz)
    one, zero = np.ones_like(eps), np.zeros_like(eps)
    C = dict(zip(("Ca", "Cb", "Da", "Db"), O0.coefficients(eps, one, zero, zero, sim.dt)))
    F = O0.zeros(nx, ny, nz, True)
    src = sim.sources[0]
    dt = sim.dt
    mons = [O0.RegionDFT({c: m._monitor_region[c] for c in m.components}, m.frequencies)
            for m in sim.monitors]
    t0 = time.perf_counter()
    O0.run(F, C, sim.grid.spacing, True, dt, steps,
           sources=[lambda F_, t: src.update_fields(F_, t, dt)], monitors=mons)
    return time.perf_counter() - t0


def cpu_sweep_baseline(steps=400, full_steps=4000, max_workers=16):
    """The sweep the reference runs on the host (ParameterSweep with a process
    pool, src/prismo/optimization/sweep.py:144-167): one process per core
    (at most 16), each running one of the 256 coupler designs (evenly
    spaced) for ``steps`` steps of the O0 reference-arithmetic step; the
    throughput is scaled to ``full_steps``-step designs.  Run before the GPU
    is touched (the workers are forked)."""
    from concurrent.futures import ProcessPoolExecutor

    import numpy as np

    workers = max(1, min(max_workers, os.cpu_count() or 1))
    ks = [int(k) for k in np.linspace(0, 255, workers).round()]
    t0 = time.perf_counter()
    with ProcessPoolExecutor(max_workers=workers) as ex:
        secs = list(ex.map(_cpu_coupler, [(k, steps) for k in ks]))
    wall = time.perf_counter() - t0
    per_design = float(np.mean(secs)) * full_steps / steps
    out = {"designs_per_s": round(workers / per_design, 4), "unit": "designs/s",
           "cores": workers, "kind": "port",
           "sample": f"{workers} designs x {steps} steps, one per process (O0 NumPy "
                     f"reference-arithmetic step, plane-wave source, 3 x 64-frequency port "
                     f"DFTs per step), scaled to {full_steps} steps; {wall:.1f} s wall"}
    print(f"[bench] CPU sweep sample: {out}", file=sys.stderr, flush=True)
    return out


def strip_bench(cpu_steps=60):
    """2-D TM strip (540 x 290 total cells, FP64, 3000 steps with an
    11-frequency line monitor): the reference-arithmetic ('compat') and the
    textbook ('yee' + CPML) step on the GPU, and the O0 NumPy restatement of
    the reference's step on one host core for a bounded number of steps."""
    import numpy as np
    import torch

    from oracle import compat as O0
    from prismo_amd import workloads as W

    out = {}
    for scheme in ("compat", "yee"):
        res = {}
        for graph in (True, False):
            sim, mon, steps = W.strip_2d(scheme=scheme)
            sim.graph = graph
            # warm-up: one 128-step block (kernels loaded, with PM_RUN_GRAPH
            # the block's graph captured)
            sim.run_steps(128)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            sim.run_steps(steps)
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            nx, ny, _ = sim.grid.dimensions
            if graph:
                res["us_per_step_graph_replay"] = round(el / steps * 1e6, 2)
            else:
                res.update({"gcells_per_s": round(nx * ny * steps / el / 1e9, 4),
                            "us_per_step": round(el / steps * 1e6, 2), "steps": steps,
                            "loop": "direct launches (the default)"})
            del sim
        out[scheme] = res
    sim, mon, _ = W.strip_2d(scheme="compat")
    nx, ny, nz = sim.grid.dimensions
    eps = W._strip_eps()["eps_rel"]
    shp = (nx, ny, nz)
    C = dict(zip(("Ca", "Cb", "Da", "Db"),
                 O0.coefficients(eps.reshape(shp), np.ones(shp), np.zeros(shp), np.zeros(shp),
                                 sim.dt)))
    F = O0.zeros(nx, ny, nz, True)
    src = sim.sources[0]
    t0 = time.perf_counter()
    O0.run(F, C, sim.grid.spacing, True, sim.dt, cpu_steps,
           sources=[lambda F_, t: src.update_fields(F_, t, sim.dt)])
    el = time.perf_counter() - t0
    out["cpu_o0"] = {"gcells_per_s": round(nx * ny * cpu_steps / el / 1e9, 5),
                     "us_per_step": round(el / cpu_steps * 1e6, 1), "steps": cpu_steps,
                     "cores": 1, "kind": "port"}
    out["cells"] = nx * ny
    out["workload"] = ("strip_2d: 540 x 290 total cells (20-cell layers), 2-D TM FP64, "
                       "500 nm Si strip, 2 um line source 193 THz, 11-frequency line monitor")
    return out


def waveguide_alg_bytes(sim, esz: int, csz: int) -> int:
    """Algorithmic bytes of one waveguide step: six field components read
    and written and three Cb per cell, two CPML memory values read and
    written per (cell, CPML axis), and the dispersive state per dispersive
    E sample (Lorentz P and Q, Drude J) read and written."""
    import numpy as np

    g = sim.grid
    n = [int(v) for v in g.dimensions]
    cells = n[0] * n[1] * n[2]
    cp = sim.solver.updater.cpml
    th = cp.thickness if cp is not None else (0,) * 6
    slab = sum((th[2 * a] + th[2 * a + 1]) * cells // n[a] for a in range(3))
    ade = 0
    for r in sim._ade:
        cnt = int(np.prod([hi - lo for lo, hi in r.box]))
        ade += cnt * (4 if r.kind == 1 else 2) * esz
    return (12 * esz + 3 * csz) * cells + 8 * esz * slab + ade


def waveguide_bench(steps=None, launch32=None, launch64=None):
    """Dispersive 3-D waveguide (480 x 240 x 160 cells, Lorentz Si core,
    Drude pad, grid-mode source, 64-frequency DFT plane): FP32 fields with
    FP32 and with FP16 Cb, against FP64 over the full run; relative L2 of
    the plane's spectra and of the flux, and the step kernel's roofline
    (algorithmic bytes per step over its HIP-event time).  launch32 /
    launch64 = (vec, ty, kchunk, order, row segments) pin the fused step's
    launch shape of the FP32 legs / the FP64 leg (for counter runs, whose
    serialised kernels would skew the tuner); by default each leg is tuned."""
    import numpy as np
    import torch

    from prismo_amd import workloads as W

    steps = steps or W.WG_STEPS
    res, spec = {}, {}
    for name, dt, storage in (("fp32", np.float32, "auto"), ("fp32_fp16cb", np.float32, "fp16"),
                              ("fp64", np.float64, "auto")):
        sim, mon, mode, _ = W.waveguide_3d(dtype=dt, coefficient_storage=storage)
        eng = sim.solver.updater.engine_for(sim.fields)
        pin = launch64 if dt == np.float64 else launch32
        if pin:
            vec, ty, kc, order, xs = pin
            eng.ctx.set_row_split(xs)
            eng.ctx.set_tuning(1, ty, kc, vec, order)
            tuned = {"vec": vec, "ty": ty, "kchunk": kc, "order": order, "xs": xs,
                     "fused": bool(eng.ctx.fused_active())}
        else:
            tuned = sim.autotune()
        sim.run_steps(2)
        sim.prepare(steps - 2)
        torch.cuda.synchronize()
        eng.ctx.set_timing(True, stride=10)
        t0 = time.perf_counter()
        sim.run_steps(steps - 2)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        f_ms, nf = eng.ctx.timing(3)
        h_ms, nh = eng.ctx.timing(0)
        e_ms, ne = eng.ctx.timing(1)
        eng.ctx.set_timing(False)
        cells = int(np.prod(W.WG_DIMS))
        flux = W.waveguide_flux(mon)
        spec[name] = {c: np.stack([np.squeeze(mon.get_frequency_data(c, f))
                                   for f in mon.frequencies]) for c in ("Ey", "Ez", "Hy", "Hz")}
        esz = 4 if dt == np.float32 else 8
        alg = waveguide_alg_bytes(sim, esz, 2 if storage == "fp16" else esz)
        k_ms = f_ms if nf >= nh else h_ms + e_ms
        res[name] = {"gcells_per_s": round(cells * (steps - 2) / el / 1e9, 3),
                     "ms_per_step": round(el / (steps - 2) * 1e3, 4), "flux": flux,
                     "fused": bool(tuned.get("fused", False)),
                     "launch": {k: tuned.get(k) for k in ("vec", "ty", "kchunk", "order", "xs")},
                     "grid_neff": float(sim.sources[0].grid_neff),
                     "roofline": {"bound": "hbm",
                                  "kernel": ("k_yee_fused (dispersive instance)" if nf >= nh
                                             else "k_yee_h + k_yee_e"),
                                  "alg_bytes_per_step": alg, "kernel_ms": round(k_ms, 5),
                                  "achieved": round(alg / (k_ms * 1e-3) / 1e9, 1),
                                  "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": round(alg / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}
        cols = sum(int(np.prod(a.shape[1:])) for a in spec[name].values())
        print(f"[bench] waveguide {name}: {el:.2f} s for {steps - 2} steps", file=sys.stderr,
              flush=True)
        del sim, mon
        torch.cuda.empty_cache()
    out = {
        "workload": "waveguide_3d: 480 x 240 x 160 cells, Lorentz Si core + Drude pad, "
                    "TE grid-mode source, 64-frequency Ey/Ez/Hy/Hz plane",
        "steps": steps, "plane_columns": cols, "frequencies": 64,
    }
    b = spec["fp64"]
    for name in ("fp32", "fp32_fp16cb"):
        a = spec[name]
        num = sum(np.sum(np.abs(a[c] - b[c]) ** 2, axis=(1, 2)) for c in a)
        den = sum(np.sum(np.abs(b[c]) ** 2, axis=(1, 2)) for c in a)
        fr = np.abs(res[name]["flux"] / res["fp64"]["flux"] - 1)
        res[name]["spectra_rel_l2_vs_fp64_max"] = float(np.sqrt(num / den).max())
        res[name]["flux_rel_err_vs_fp64_max"] = float(fr.max())
    for name in res:
        out[name] = {k: v for k, v in res[name].items() if k != "flux"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--launch", type=str, default=None)
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--batch-probe", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--surrogate", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--sweep-only", action="store_true")
    ap.add_argument("--no-strip", action="store_true")
    ap.add_argument("--no-waveguide", action="store_true")
    ap.add_argument("--waveguide-only", action="store_true")
    ap.add_argument("--waveguide-steps", type=int, default=None)
    ap.add_argument("--waveguide-launch", type=str, default=None,
                    help="vec,ty,kchunk,order,xs of the FP32 waveguide legs (default: tuned)")
    ap.add_argument("--waveguide-launch64", type=str, default=None,
                    help="vec,ty,kchunk,order,xs of the FP64 waveguide leg (default: tuned)")
    ap.add_argument("--fp64-only", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only; runs on several GPUs are not supported "
              "(the 3-D stencil would shard by z-slabs)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    if args.batch_probe:
        batch_probe()
        return

    # the host-process sweep baseline first, while no process has touched
    # the GPU (its workers are forked from this one)
    runs_sweep = not (args.no_sweep or args.waveguide_only or args.surrogate)
    cpu_sweep = (cpu_sweep_baseline() if runs_sweep and not args.no_cpu_baseline
                 else None)

    import numpy as np
    import torch

    import prismo_amd as pa
    from prismo_amd.workloads import voxel_3d

    pa.set_backend("mi355x", 0)
    if args.sweep_only:
        sweep = sweep_bench()
        if cpu_sweep is not None:
            sweep["cpu_baseline"] = cpu_sweep
            sweep["speedup_vs_cpu_pool"] = round(sweep["designs_per_s"]
                                                 / cpu_sweep["designs_per_s"], 1)
        print(json.dumps({"sweep": sweep}), flush=True)
        return
    pin = lambda t: tuple(int(x) for x in t.split(",")) if t else None
    if args.waveguide_only:
        print(json.dumps({"waveguide_3d": waveguide_bench(args.waveguide_steps,
                                                          pin(args.waveguide_launch),
                                                          pin(args.waveguide_launch64))}),
              flush=True)
        return
    if args.fp64_only:
        print(json.dumps({"voxel_fp64": voxel_fp64(args.n, args.steps, args.warmup, None)}),
              flush=True)
        return
    if args.surrogate:
        print(json.dumps({"surrogate": surrogate_bench()}), flush=True)
        return
    sim = voxel_3d(dtype=np.float32, coefficient_storage="fp16", n=args.n)
    eng = sim.solver.updater.engine_for(sim.fields)
    print(f"[bench] built {args.n}^3 FP32/FP16 voxel workload", file=sys.stderr, flush=True)
    if args.launch:
        parts = [int(x) for x in args.launch.split(",")]
        vec, ty, kc, order = parts[:4]
        nstep = parts[4] if len(parts) > 4 else 1
        xs = parts[5] if len(parts) > 5 else 1
        eng.ctx.set_row_split(xs)
        eng.ctx.set_tuning(nstep, ty, kc, vec, order)
        tuned = {"steps": nstep, "vec": vec, "ty": ty, "kchunk": kc, "order": order,
                 "xs": xs, "fused": True}
    else:
        tuned = sim.autotune()
    print(f"[bench] fused step launch: {tuned}", file=sys.stderr, flush=True)
    sim.run_steps(args.warmup)
    # the source tables of the timed steps go to the device before the clock
    # starts
    sim.prepare(args.steps)
    torch.cuda.synchronize()
    # HIP events around the kernels of every 10th step (every step of a
    # short run): enough launches for a stable average, few enough that the
    # events do not slow the run
    eng.ctx.set_timing(True, stride=max(1, min(10, args.steps // 20)))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.run_steps(args.steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    h_ms, nh = eng.ctx.timing(0)
    e_ms, ne = eng.ctx.timing(1)
    f_ms, nf = eng.ctx.timing(3)
    f2_ms, nf2 = eng.ctx.timing(5)
    eng.ctx.set_timing(False)
    cells = args.n**3
    value = cells * args.steps / el / 1e9
    # the kernel that advances (nearly) every step: the two-step pass (two
    # launches, slab tiles and interior tiles), else the fused H+E kernel,
    # else the separate H and E kernels
    per_step_bytes = algorithmic_bytes(eng.plan)
    if nf2 and nf2 >= nf:
        kern_ms, kname, steps_per_launch = (
            f2_ms, "k_yee_fused2_slab + k_yee_fused2_in (two steps per pass)", 2)
    elif nf >= nh:
        kern_ms, kname, steps_per_launch = f_ms, "k_yee_fused (H+E, one pass per step)", 1
    else:
        kern_ms, kname, steps_per_launch = h_ms + e_ms, "k_yee_h + k_yee_e (one step)", 1
    alg_bytes = per_step_bytes * steps_per_launch
    achieved = alg_bytes / (kern_ms * 1e-3) / 1e9
    launch = {k: tuned.get(k) for k in ("steps", "vec", "ty", "kchunk", "order", "xs")}
    wkey = f"voxel_3d_{args.n}_fp32_fp16"
    traffic = measured_traffic(wkey, launch)
    finite = all(bool(torch.isfinite(sim.fields[c]).all()) for c in ("Ex", "Ey", "Ez"))
    ref = {c: sim.fields[c].clone() for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")}
    print(f"[bench] {args.steps} steps in {el:.3f} s, two-step {f2_ms:.4f} ms x{nf2}, "
          f"fused {f_ms:.4f} ms x{nf}, H {h_ms:.4f} ms x{nh}, E {e_ms:.4f} ms x{ne}, "
          f"finite={finite}", file=sys.stderr, flush=True)
    del sim, eng
    torch.cuda.empty_cache()
    accuracy = None if args.no_fp64 else voxel_accuracy(args.n, tuned)
    fp64 = None if args.no_fp64 else voxel_fp64(args.n, args.steps, args.warmup, ref)
    del ref
    torch.cuda.empty_cache()
    sweep = None if args.no_sweep else sweep_bench()
    if sweep is not None and cpu_sweep is not None:
        sweep["cpu_baseline"] = cpu_sweep
        sweep["speedup_vs_cpu_pool"] = round(sweep["designs_per_s"] / cpu_sweep["designs_per_s"], 1)
    strip = None if args.no_strip else strip_bench()
    surrogate = None if args.no_surrogate else surrogate_bench()
    waveguide = None if args.no_waveguide else waveguide_bench()
    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] timing the CPU oracle sample", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.n, args.cpu_steps)
    line = {
        "metric": "Gcell-updates/s (3-D Yee+CPML, FP32 fields / FP16 coefficients)",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"voxel_3d: {args.n}^3 Yee+CPML(12) at 20 nm, eps voxels 8^3 "
                        "{1.0,2.085,12.1} seed 20261202, Ez dipole 193 THz",
            "cells": cells,
            "coefficients": "fp16 Cb",
            "parallelism": "single GPU",
            "launch": launch,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kname,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "alg_bytes_per_launch": alg_bytes,
            "steps_per_launch": steps_per_launch,
            "kernel_ms": round(kern_ms, 5),
        },
        "cpu_baseline": cpu,
        "voxel_accuracy": accuracy,
        "voxel_fp64": fp64,
        "sweep": sweep,
        "strip_2d": strip,
        "waveguide_3d": waveguide,
        "surrogate": surrogate,
    }
    print(json.dumps(line), flush=True)


def surrogate_bench() -> dict:
    """The surrogate loop end to end: 4096 coupler designs (gap, length,
    width, cladding) through the batched FDTD for training spectra, a
    4-member BF16 MLP ensemble trained on the matrix cores, 65536 designs
    screened, the top 64 plus the uncertain ones re-run by the batched FDTD
    as judge.  ``accepted_per_s`` counts only designs the FDTD confirmed."""
    from prismo_amd.surrogate import loop

    return loop.run(n_train=4096, n_screen=65536, top_k=64, epochs=300,
                    log=lambda m: print(m, file=sys.stderr, flush=True))


if __name__ == "__main__":
    main()
