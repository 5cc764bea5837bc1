"""High-level Simulation (src/prismo/core/simulation.py:22-250).

Loop semantics are the reference's: each step is H update, E update,
``t += dt`` accumulated in FP64, every source at the new time, then every
monitor (simulation.py:147-164).  On the mi355x backend whole runs go to the
step library: times are accumulated on the host exactly as the reference
does, each source becomes per-step value tables, each monitor a sampled box
with its spectra summed block-wise on the device.  Sources or monitors
without a device form are still honoured, by stepping one at a time and
calling them on the device field views.

Additions for the configurations this package targets (none exist in the
reference): ``material_arrays=`` (as ``MaxwellUpdater`` takes),
``scheme=``/``pml=`` (textbook stencil, CPML) and ``add_dispersive_box``.
"""

from __future__ import annotations

import time as time_module
from typing import Optional

import numpy as np

from .. import _native as N
from ..backends.backend_manager import resolve_backend
from .engine import AdeRegion
from .fields import ElectromagneticFields
from .grid import GridSpec, YeeGrid
from .solver import FDTDSolver


class Simulation:
    def __init__(self, size, resolution, boundary_conditions: str = "pml", pml_layers: int = 10,
                 courant_factor: float = 0.9, *, backend=None, scheme: str = "compat",
                 pml=None, material_arrays: Optional[dict] = None, dtype=None,
                 coefficient_storage: str = "auto"):
        self.grid_spec = GridSpec(size=size, resolution=resolution, boundary_layers=pml_layers)
        self.grid = YeeGrid(self.grid_spec)
        self.size = size
        self.resolution = resolution
        self.boundary_conditions = boundary_conditions
        self.courant_factor = courant_factor
        self.backend = resolve_backend(backend)
        self.fields = ElectromagneticFields(self.grid, dtype=dtype, backend=self.backend)
        self.dt = self.grid.get_time_step(courant_factor)
        self._solver_kw = dict(scheme=scheme, pml=pml, coefficient_storage=coefficient_storage)
        self._material_arrays = material_arrays
        self._ade: list = []
        self.solver = FDTDSolver(self.grid, self.dt, material_arrays, backend=self.backend,
                                 fields=self.fields, **self._solver_kw)
        self.sources: list = []
        self.monitors: list = []
        self.step_count = 0
        self.current_time = 0.0
        # replay blocks of steps as captured HIP graphs (True) or launch
        # every step's kernels directly (False, the default: on the 540 x 290
        # strip both take the same time, 12.5 us per textbook step, because
        # the device time of its four dependent kernels, not the host's
        # launches, sets the pace); the results are the same bits either way
        self.graph: Optional[bool] = False

    # -- set-up -------------------------------------------------------------------
    def add_source(self, source) -> None:
        if hasattr(source, "bind_simulation"):
            source.bind_simulation(self)
        source.initialize(self.grid)
        self.sources.append(source)

    def add_monitor(self, monitor) -> None:
        monitor.initialize(self.grid)
        self.monitors.append(monitor)

    def add_structure(self, shape) -> None:
        """Paint ``shape`` (prismo_amd.geometry) over the structures added
        before it; the grid is rasterised and the coefficients built on the
        device when the simulation first steps.  Lossless, non-magnetic
        materials; the background is ``set_background`` (vacuum by default)."""
        self._structures = getattr(self, "_structures", []) + [shape]
        self.solver.updater.set_structures(self._structures, getattr(self, "_background", 1.0))

    def set_background(self, eps_rel: float) -> None:
        """Relative permittivity of the nodes no structure covers."""
        self._background = float(getattr(eps_rel, "epsilon_r", eps_rel))
        if getattr(self, "_structures", None):
            self.solver.updater.set_structures(self._structures, self._background)

    def add_dispersive_box(self, material, box_min, box_max) -> list:
        """Make every E sample point inside [box_min, box_max] dispersive
        (Lorentz with one pole, or Drude); needs ``scheme='yee'``.  Returns
        the per-component index boxes."""
        from ..materials.dispersion import DrudeMaterial, LorentzMaterial

        if self._solver_kw["scheme"] != "yee":
            raise NotImplementedError("dispersive regions need scheme='yee'")
        co = material.get_ade_coefficients(self.dt)
        if isinstance(material, LorentzMaterial):
            if len(co["poles"]) != 1:
                raise NotImplementedError("one Lorentz pole per region")
            p = co["poles"][0]
            kind, c = 1, (p["C0"], p["C1"], p["C2"], p["C3"])
        elif isinstance(material, DrudeMaterial):
            kind, c = 2, (co["C0"], co["C1"], 0.0, 0.0)
        else:
            raise NotImplementedError(f"{type(material).__name__} has no coupled ADE form")
        comps = ["Ez"] if self.grid.is_2d else ["Ex", "Ey", "Ez"]
        boxes = []
        for comp in comps:
            box = self._inside_box(comp, box_min, box_max)
            if box is None:
                continue
            self._ade.append(AdeRegion(comp, box, kind, float(material.epsilon_inf), c))
            boxes.append((comp, box))
        self.solver.updater.ade = list(self._ade)
        self.solver.updater.invalidate()
        return boxes

    def _inside_box(self, comp, lo, hi):
        coords = self.grid.get_coordinates(comp)
        rng = []
        for ax in range(len(coords)):
            sel = np.nonzero((coords[ax] >= lo[ax]) & (coords[ax] <= hi[ax]))[0]
            if sel.size == 0:
                return None
            rng.append((int(sel[0]), int(sel[-1]) + 1))
        if self.grid.is_2d:
            rng.append((0, 1))
        return tuple(rng)

    def autotune(self, force: bool = False) -> dict:
        """Pick the fastest launch shape of the fused step for this grid
        (state is restored afterwards; see ``prismo_amd.autotune``)."""
        from ..autotune import tune

        return tune(self.solver.updater.engine_for(self.fields), force=force)

    # -- running --------------------------------------------------------------------
    def run(self, time: float, progress_callback=None, progress_interval: int = 10) -> None:
        steps = int(np.ceil(time / self.dt))
        start = time_module.time()
        self.run_steps(steps, progress_callback, progress_interval, _start=start)
        if progress_callback is not None:
            progress_callback(steps, steps, self.current_time, time_module.time() - start)

    def step(self) -> None:
        self.run_steps(1)

    def _io_plan(self, steps: int):
        """Per-step tables of the next ``steps`` steps: times accumulated in
        FP64 from the current time, device source terms, monitors."""
        eng = self.solver.updater.engine_for(self.fields)
        dt = self.dt
        times = []
        t = self.current_time
        for _ in range(steps):
            t += dt
            times.append(t)
        times_arr = np.array(times)
        terms, py_sources = [], []
        for s in self.sources:
            tt = s.device_terms(times)
            if tt is None:
                py_sources.append(s)
            else:
                terms.extend(tt)
        dev_mons, py_mons = [], []
        for m in self.monitors:
            (py_mons if m.device_spec() is None else dev_mons).append(m)
        esz = 8 if eng.precision == N.FP64 else 4
        ncols = sum(int(np.prod(sh)) for m in dev_mons for _, _, sh in m.device_spec()["boxes"])
        any_td = any(m.device_spec()["time_domain"] for m in dev_mons)
        # rows per spectral block: FP32 sums each block on the matrix cores
        # before the compensated add, so its blocks stay short; a small grid
        # replays blocks as graphs, whose length is the block's
        T = 128 if esz == 4 or self._small_grid() else 512
        if any_td and ncols:
            T = int(max(1, min(T, steps, (256 << 20) // max(1, ncols * esz))))
        return eng, times_arr, terms, py_sources, dev_mons, py_mons, any_td, T

    def prepare(self, steps: int) -> None:
        """Upload the source and monitor tables of the next ``steps`` steps
        now, so that a following ``run_steps(steps)`` starts stepping at
        once (it uses these tables as long as nothing was stepped, added or
        prepared in between)."""
        if steps <= 0:
            return
        plan = self._io_plan(steps)
        eng, times_arr, terms, _, dev_mons, _, _, T = plan
        eng.configure_io(terms, dev_mons, times_arr, step0=self.step_count, sample_T=T)
        self._prepared = ((self.step_count, steps, len(self.sources), len(self.monitors),
                           id(eng)), plan)

    def run_steps(self, steps: int, progress_callback=None, progress_interval: int = 10,
                  _start=None) -> None:
        if steps <= 0:
            return
        start = time_module.time() if _start is None else _start
        dt = self.dt
        step0 = self.step_count
        prep = getattr(self, "_prepared", None)
        self._prepared = None
        eng = self.solver.updater.engine_for(self.fields)
        key = (step0, steps, len(self.sources), len(self.monitors), id(eng))
        if prep is not None and prep[0] == key:
            eng, times_arr, terms, py_sources, dev_mons, py_mons, any_td, T = prep[1]
        else:
            eng, times_arr, terms, py_sources, dev_mons, py_mons, any_td, T = self._io_plan(steps)
            eng.configure_io(terms, dev_mons, times_arr, step0=step0, sample_T=T)
        for m in dev_mons:
            if hasattr(m, "_note_dt"):
                m._note_dt(dt)
        flags = N.RUN_SOURCES | N.RUN_SAMPLE | N.RUN_DFT
        per_step = bool(py_sources or py_mons)
        if self._use_graph(eng, T):
            flags |= N.RUN_GRAPH
        n = step0
        end = step0 + steps
        while n < end:
            # time-domain samples come back one block of rows at a time; the
            # spectra are summed by the library at block ends within any run
            seg = min(T - (n % T), end - n) if any_td else end - n
            if per_step:
                seg = 1
            if progress_callback is not None:
                # stop after every step i (0-based in this run) with i % interval == 0
                i = n - step0
                nxt = i if i % progress_interval == 0 else i + (progress_interval - i % progress_interval)
                seg = min(seg, nxt - i + 1)
            seg_times = times_arr[n - step0:n - step0 + seg]
            if py_sources:
                # every source before any monitor (simulation.py:158-164):
                # step with the device sources, add the host ones, then
                # sample the device monitors from those fields
                eng.run(n, seg, N.RUN_SOURCES)
                for s in py_sources:
                    s.update_fields(self.fields, float(seg_times[-1]), dt)
                if dev_mons:
                    eng.run(n, 1, N.RUN_SAMPLE | N.RUN_DFT | N.RUN_IO_ONLY)
            else:
                eng.run(n, seg, flags)
            if any_td:
                eng.sync()
                rows = eng.sample_rows(n % T, seg)
            for b in eng.mon:
                per = {}
                if b.spec["time_domain"]:
                    per = {c: rows[:, c0:c1] for c, (c0, c1) in b.cols.items()}
                b.monitor._ingest_samples(seg_times, per)
            self.step_count = n + seg
            self.current_time = float(seg_times[-1])
            self.solver.time = self.current_time
            self.solver.step_count = self.step_count
            if per_step:
                for m in py_mons:
                    m.update(self.fields, self.current_time, dt)
            if progress_callback is not None:
                i = self.step_count - 1 - step0
                if i % progress_interval == 0:
                    progress_callback(i, steps, self.current_time, time_module.time() - start)
            n += seg
        eng.flush()
        for b in eng.mon:
            if b.acc is not None:
                b.monitor._ingest_spectra(eng.spectra(b))

    def _use_graph(self, eng, T: int) -> bool:
        if not self.graph or T % 2:
            return False
        # two steps per launch (large grids) are not captured
        return eng.ctx.fused_steps() != 2

    def _small_grid(self) -> bool:
        """Launch-bound: 2-D, or 3-D up to 2^21 cells."""
        g = self.grid.dimensions
        return self.grid.is_2d or int(g[0]) * int(g[1]) * int(g[2]) <= (1 << 21)

    # -- results --------------------------------------------------------------------
    def get_field_data(self, monitor, component):
        if monitor not in self.monitors:
            raise ValueError("Monitor not found in this simulation")
        return monitor.get_time_data(component)[1]

    def get_frequency_data(self, monitor, component, frequency):
        if monitor not in self.monitors:
            raise ValueError("Monitor not found in this simulation")
        return monitor.get_frequency_data(component, frequency)

    def get_transmission(self, monitor, frequency=None) -> float:
        if monitor not in self.monitors:
            raise ValueError("Monitor not found in this simulation")
        if frequency is None:
            _, pf = monitor.get_power_flow()
        else:
            pf = monitor.get_power_flow(frequency)
        return np.mean(pf)
