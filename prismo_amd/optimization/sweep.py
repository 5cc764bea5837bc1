"""Parameter sweeps (API of src/prismo/optimization/sweep.py:18-300).

``SweepParameter`` / ``ParameterSweep`` keep the reference's interface: the
combinations are the meshgrid of the parameter values (sweep.py:85-103),
``run`` returns one dict per combination with ``parameters`` and
``status`` ('success' or 'error' with the message, sweep.py:169-195),
``save_results`` / ``get_result_array`` / ``find_optimal`` as there.

Added: ``ParameterSweep.from_samples`` for a list of parameter dicts (random
design sets), and a batched back end.  With ``batched=BatchedRunner(...)``
the sweep builds every design's Simulation with ``build``, advances them
together on the GPU (``prismo_amd.core.batch.BatchEngine``, one launch per
step over every design of a group of ``max_batch``), then calls
``analyze(sim, params)`` per design.  Groups are pipelined: group g+1 is
built and queued on its own stream while group g runs, and group g is
analysed while group g+1 runs.  A design whose build or analysis raises is
reported with status 'error' and does not stop the others; so is one whose
fields leave the finite numbers (the kernel flags each design that computes
an Inf or NaN).  A design the batch cannot take (3-D, lossy, a time-domain
monitor, another source waveform or frequency list than the group's) runs
alone through its own ``Simulation.run_steps``.
"""

from __future__ import annotations

import json
from collections import Counter
from concurrent.futures import ProcessPoolExecutor
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, Callable, Optional

import numpy as np


@dataclass
class SweepParameter:
    name: str
    values: list
    unit: str = ""

    def __len__(self) -> int:
        return len(self.values)


@dataclass
class BatchedRunner:
    """build(params) -> Simulation (not run); analyze(sim, params) -> dict.
    Designs run in groups of ``max_batch``; ``first_batch`` (if set) sizes
    the first group alone, so the GPU starts after a shorter build."""
    build: Callable[[dict], Any]
    analyze: Callable[[Any, dict], dict]
    steps: int
    dtype: Any = np.float32
    coefficient_storage: str = "auto"
    max_batch: int = 256
    graph: bool = True
    engine_kw: dict = field(default_factory=dict)
    first_batch: Optional[int] = None


class ParameterSweep:
    def __init__(self, parameters: list, simulation_func: Optional[Callable] = None,
                 output_dir: Optional[Path] = None, parallel: bool = False,
                 num_workers: Optional[int] = None, batched: Optional[BatchedRunner] = None):
        self.parameters = list(parameters)
        self.simulation_func = simulation_func
        self.output_dir = Path(output_dir) if output_dir else Path("./sweep_results")
        self.parallel = parallel
        self.num_workers = num_workers
        self.batched = batched
        self.results: list = []
        self.parameter_combinations: list = []
        self._generate_combinations()

    @classmethod
    def from_samples(cls, samples: list, simulation_func: Optional[Callable] = None, **kw):
        sw = cls([], simulation_func, **kw)
        sw.parameter_combinations = [dict(s) for s in samples]
        return sw

    def _generate_combinations(self) -> None:
        if not self.parameters:
            return
        grids = np.meshgrid(*[p.values for p in self.parameters], indexing="ij")
        for i in range(grids[0].size):
            self.parameter_combinations.append(
                {p.name: grids[k].flat[i] for k, p in enumerate(self.parameters)})

    # ----------------------------------------------------------------- running
    def run(self, show_progress: bool = False) -> list:
        if self.batched is not None:
            results = self._run_batched()
        elif self.parallel:
            results = self._run_parallel()
        else:
            results = [self._run_single_simulation(p) for p in self.parameter_combinations]
        self.results = results
        return results

    def _run_parallel(self) -> list:
        with ProcessPoolExecutor(max_workers=self.num_workers) as ex:
            futs = [ex.submit(self._run_single_simulation, p) for p in self.parameter_combinations]
            return [f.result() for f in futs]

    def _run_single_simulation(self, params: dict) -> dict:
        try:
            result = self.simulation_func(params)
            result["parameters"] = params
            result["status"] = "success"
            return result
        except Exception as e:  # per-design failure, as sweep.py:190-195
            return {"parameters": params, "status": "error", "error": str(e)}

    def _run_batched(self) -> list:
        from ..core.batch import BatchEngine

        br = self.batched
        combos = self.parameter_combinations
        results: list = [None] * len(combos)

        def fail(keys, e):
            for k in keys:
                results[k] = {"parameters": combos[k], "status": "error", "error": str(e)}

        def complete(eng, group):
            try:
                eng.finish()
            except Exception as e:
                fail([k for k, _ in group], e)
                return
            for d, (k, sim) in enumerate(group):
                if eng.diverged[d]:
                    # the reference would report NaN results as a success;
                    # a run that left the finite numbers is a failed design
                    fail([k], "fields diverged (Inf or NaN) during the batched run")
                    continue
                try:
                    r = dict(br.analyze(sim, combos[k]))
                    r["parameters"] = combos[k]
                    r["status"] = "success"
                    results[k] = r
                except Exception as e:
                    fail([k], e)

        def run_alone(k, sim):
            try:
                sim.run_steps(br.steps)
                r = dict(br.analyze(sim, combos[k]))
                r["parameters"] = combos[k]
                r["status"] = "success"
                results[k] = r
            except Exception as e:
                fail([k], e)

        pending = None  # (engine, group) queued on the GPU, not yet analysed
        starts = [0]
        first = br.first_batch if br.first_batch else br.max_batch
        starts += list(range(min(first, len(combos)), len(combos), br.max_batch))
        for gi, g0 in enumerate(starts):
            g1 = starts[gi + 1] if gi + 1 < len(starts) else len(combos)
            keyed, singles = [], []
            for k in range(g0, g1):
                try:
                    sim = br.build(combos[k])
                except Exception as e:
                    fail([k], e)
                    continue
                try:
                    keyed.append((k, sim, BatchEngine.design_key(sim, br.steps)))
                except NotImplementedError:
                    singles.append((k, sim))
            # the group runs the designs sharing the most common key; a
            # design a batch cannot take, or one of another key, runs alone
            group = []
            if keyed:
                main = Counter(key for _, _, key in keyed).most_common(1)[0][0]
                group = [(k, sim) for k, sim, key in keyed if key == main]
                singles += [(k, sim) for k, sim, key in keyed if key != main]
            eng = None
            if group:
                try:
                    eng = BatchEngine([s for _, s in group], dtype=br.dtype,
                                      coefficient_storage=br.coefficient_storage, **br.engine_kw)
                    eng.start(br.steps, graph=br.graph)
                except Exception as e:
                    fail([k for k, _ in group], e)
                    eng = None
            if pending is not None:
                complete(*pending)
            pending = (eng, group) if eng is not None else None
            for k, sim in sorted(singles):
                run_alone(k, sim)
        if pending is not None:
            complete(*pending)
        return results

    # ----------------------------------------------------------------- results
    def save_results(self, filename: str = "sweep_results.json") -> Path:
        self.output_dir.mkdir(parents=True, exist_ok=True)
        path = self.output_dir / filename

        def conv(v):
            if isinstance(v, np.ndarray):
                if np.iscomplexobj(v):
                    return {"real": v.real.tolist(), "imag": v.imag.tolist()}
                return v.tolist()
            if isinstance(v, (np.floating, np.integer)):
                return v.item()
            if isinstance(v, dict):
                return {k: conv(x) for k, x in v.items()}
            return v

        with open(path, "w") as f:
            json.dump([{k: conv(v) for k, v in r.items()} for r in self.results], f, indent=2)
        return path

    def get_result_array(self, result_key: str) -> np.ndarray:
        values = [r.get(result_key, np.nan) for r in self.results]
        if self.parameters:
            return np.array(values).reshape(tuple(len(p.values) for p in self.parameters))
        return np.array(values)

    def find_optimal(self, metric: str, maximize: bool = True):
        ok = [r for r in self.results if r.get("status") == "success" and metric in r]
        if not ok:
            raise ValueError(f"No successful results with metric '{metric}'")
        best = (max if maximize else min)(ok, key=lambda r: r[metric])
        return best["parameters"], best
