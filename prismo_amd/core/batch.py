"""Batched 2-D TM stepping of many Simulations at once (include/pmbatch.h).

A parameter sweep in the reference runs one ``Simulation`` per parameter set
(src/prismo/optimization/sweep.py:119-167), each its own Python step loop.
``BatchEngine`` takes the Simulations of a sweep as descriptions -- grid,
materials, CPML, sources, monitors, exactly as they were built -- and
advances all of them together: one kernel launch per step over every design,
the step loop replayed as a captured HIP graph, the spectra of every port of
every design one MFMA GEMM per block of steps.  Results land in each
Simulation's own monitors, so analysis code reads them as after
``Simulation.run``.

What a batch takes: 2-D, ``scheme='yee'``, lossless non-magnetic media, a
``PMLParams`` CPML, sources whose device form is an Ez and/or Hy line along
y with the same waveform in every design (``PlaneWaveSource``, line
``PointSource``), and ``FieldMonitor`` lines along y (Ez, Hx, Hy) with one
shared frequency list.  Anything else raises ``NotImplementedError``.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _native as N
from ..monitors.field import FieldMonitor
from .coefficients import EPS0, component_average, is_lossless_nonmagnetic

_COMP_LINE = {"Ez": 0, "Hx": 1, "Hy": 2}


def _torch():
    import torch

    return torch


class BatchEngine:
    def __init__(self, sims, dtype=np.float32, coefficient_storage: str = "auto",
                 chunk=8, sample_T: int = 128, device: int = 0):
        torch = _torch()
        if not sims:
            raise ValueError("empty batch")
        self.sims = list(sims)
        self.precision = N.FP64 if np.dtype(dtype) == np.float64 else N.FP32
        self.tdtype = torch.float64 if self.precision == N.FP64 else torch.float32
        self.device = torch.device("cuda", device)
        self.stream = torch.cuda.Stream(device=self.device)  # default priority
        if coefficient_storage == "fp16":
            if self.precision != N.FP32:
                raise ValueError("FP16 coefficients need FP32 fields")
            self.coeff_kind = N.COEFF_CB_F16
        elif coefficient_storage in ("auto", "native"):
            self.coeff_kind = N.COEFF_CB
        else:
            raise ValueError("coefficient_storage must be 'auto', 'native' or 'fp16'")
        s0 = self.sims[0]
        self.dt = float(s0.dt)
        self.d = tuple(s0.grid.spacing)
        for s in self.sims:
            if not s.grid.is_2d:
                raise NotImplementedError("batched stepping is 2-D")
            if s.solver.updater.scheme != "yee":
                raise NotImplementedError("batched stepping needs scheme='yee'")
            if s.dt != self.dt or tuple(s.grid.spacing) != self.d:
                raise NotImplementedError("every design must share dx, dy and dt")
        self.sample_T = int(sample_T)
        # rows per tile; "auto": the autotuner picks them with the launch
        # shape when the run starts (prismo_amd.autotune.tune_batch)
        self.auto_tune = chunk == "auto"
        self.chunk = 8 if self.auto_tune else int(chunk)
        self.tuning = None
        self._keep: list = []
        self.step_count = 0
        self._cb_q: list = []   # per design: (stored Cb, FP16 scale or None)
        self._times_done = 0.0
        self.ctx = N.BatchContext(device, self.stream.cuda_stream)
        self._built = False

    @property
    def cb_host(self) -> list:
        """Per design, the Ez-point Cb the kernel applies, FP64 (the FP16
        values times their scale for FP16 storage)."""
        return [c if sc is None else c.astype(np.float64) * sc for c, sc in self._cb_q]

    # ------------------------------------------------------------------ setup
    def _alloc(self, shape, dtype=None):
        t = _torch().zeros(shape, dtype=dtype or self.tdtype, device=self.device)
        self._keep.append(t)
        return t

    _FIELD_ARRAYS = ("ez", "hx", "hy", "psi_hx", "psi_hy")

    def _design(self, s, des, col):
        """Geometry, coefficients, CPML and sample lines of one design; the
        device arrays are placed afterwards in shared arenas."""
        up = s.solver.updater
        g = s.grid
        nx, ny, _ = g.dimensions
        sy = ((nx + 63) // 64) * 64
        des.nx, des.ny, des.sy = nx, ny, sy
        mu, se, sm = up.mu_rel, up.sigma_e, up.sigma_m
        if not is_lossless_nonmagnetic(mu, se, sm):
            raise NotImplementedError("batched stepping needs lossless, non-magnetic media")
        Db = np.asarray(up.Db)
        des.db = float(Db.flat[0])
        cb = component_average(np.asarray(up.Cb)[:, :, 0], "Ez", True)   # (nx-1, ny-1)
        scale = self.dt / EPS0
        if self.coeff_kind == N.COEFF_CB_F16:
            q = (cb / scale).astype(np.float16)
            cb_pad = np.zeros((ny, sy), np.float16)
            cb_pad[: ny - 1, : nx - 1] = q.T
            des.cb_scale = scale
            self._cb_q.append((q, scale))
        else:
            cb_pad = np.zeros((ny, sy), np.float64 if self.precision == N.FP64 else np.float32)
            cb_pad[: ny - 1, : nx - 1] = cb.T
            des.cb_scale = 1.0
            self._cb_q.append((cb, None))
        # CPML profiles: x as six arrays of nx (be ce ike bh ch ikh), then y
        # interleaved per row, 8 values (the same six, two unused)
        cp = up.cpml
        th = cp.thickness if cp is not None else (0, 0, 0, 0, 0, 0)
        des.pml[0], des.pml[1], des.pml[2], des.pml[3] = th[0], th[1], th[2], th[3]
        keys = ("be", "ce", "ike", "bh", "ch", "ikh")
        ident = (1.0, 0.0, 1.0, 1.0, 0.0, 1.0)
        prof = np.zeros(6 * nx + 8 * ny)
        px = prof[:6 * nx].reshape(6, nx)
        py = prof[6 * nx:].reshape(ny, 8)
        for k, v in enumerate(ident):
            px[k] = v
            py[:, k] = v
        if cp is not None:
            pr = cp.profiles(self.dt)
            for k, key in enumerate(keys):
                if pr[0] is not None:
                    px[k] = pr[0][key]
                if pr[1] is not None:
                    py[:, k] = pr[1][key]
        self._host_cb.append(cb_pad)
        self._host_prof.append(prof)
        # monitors -> sample lines
        nl = 0
        mons = []
        for m, lines in self._monitor_lines(s):
            cols = {}
            for comp, i0, j0, j1 in lines:
                L = des.line[nl]
                L.comp, L.i, L.j0, L.j1, L.col = _COMP_LINE[comp], i0, j0, j1, col
                cols[comp] = (col, col + (j1 - j0))
                col += j1 - j0
                nl += 1
            mons.append((m, cols))
        des.n_lines = nl
        return col, mons

    @staticmethod
    def _monitor_lines(s) -> list:
        """Per monitor of design s, its sample lines (comp, i, j0, j1)."""
        out = []
        nl = 0
        for m in s.monitors:
            if not isinstance(m, FieldMonitor) or m.time_domain or not m.frequencies:
                raise NotImplementedError("batched designs take frequency-domain FieldMonitors")
            lines = []
            for comp, box, shape in m.device_spec()["boxes"]:
                (i0, i1), (j0, j1), _ = box
                if comp not in _COMP_LINE or i1 - i0 != 1:
                    raise NotImplementedError("batched monitors are Ez/Hx/Hy lines along y")
                if nl >= N.BATCH_MAX_LINES:
                    raise NotImplementedError("at most 8 sample lines per design")
                lines.append((comp, i0, j0, j1))
                nl += 1
            out.append((m, lines))
        return out

    @classmethod
    def design_key(cls, s, nsteps: int) -> tuple:
        """For a Simulation a batch can take, what every design of one batch
        must share (dt, dx and dy, the monitors' frequency list, the source
        waveforms over ``nsteps`` steps); raises ``NotImplementedError`` with
        the reason for one it cannot take.  A sweep groups designs by this
        key and runs the others one at a time."""
        if not s.grid.is_2d:
            raise NotImplementedError("batched stepping is 2-D")
        up = s.solver.updater
        if up.scheme != "yee":
            raise NotImplementedError("batched stepping needs scheme='yee'")
        if not is_lossless_nonmagnetic(up.mu_rel, up.sigma_e, up.sigma_m):
            raise NotImplementedError("batched stepping needs lossless, non-magnetic media")
        freqs = None
        for m, _ in cls._monitor_lines(s):
            f = tuple(float(v) for v in m.frequencies)
            if freqs is None:
                freqs = f
            elif f != freqs:
                raise NotImplementedError("every monitor of a batch shares one frequency list")
        times = []
        t = 0.0
        for _ in range(nsteps):
            t += s.dt
            times.append(t)
        e, h = cls._sources(s, np.array(times))
        return (float(s.dt), tuple(float(v) for v in s.grid.spacing), freqs,
                None if e is None else e[3].tobytes(), None if h is None else h[3].tobytes())

    @staticmethod
    def _sources(s, times):
        e = h = None
        for src in s.sources:
            terms = src.device_terms(times)
            if terms is None:
                raise NotImplementedError(f"{type(src).__name__} has no device form")
            for t in terms:
                (i0, i1), (j0, j1), _ = t.box
                if t.profile is not None or t.divisor != 0.0 or i1 - i0 != 1:
                    raise NotImplementedError("batched sources are uniform lines along y")
                if t.comp == "Ez" and e is None:
                    e = (i0, j0, j1, np.asarray(t.table, dtype=np.float64))
                elif t.comp == "Hy" and h is None:
                    h = (i0, j0, j1, np.asarray(t.table, dtype=np.float64))
                else:
                    raise NotImplementedError("one Ez line and one Hy line per design")
        return e, h

    def _place_arrays(self, arr):
        """One zero-filled arena for every design's fields and CPML memory,
        one upload each for the coefficients and the profiles."""
        torch = _torch()
        n = len(self.sims)
        sizes = [arr[d].ny * arr[d].sy for d in range(n)]
        per = 2 * len(self._FIELD_ARRAYS) + 2
        arena = self._alloc((per * sum(sizes),))
        self._arena = arena
        cb_all = torch.from_numpy(np.concatenate([c.ravel() for c in self._host_cb]))
        cbd = cb_all.to(self.device, torch.float16 if self.coeff_kind == N.COEFF_CB_F16
                        else self.tdtype)
        pr_all = torch.from_numpy(np.concatenate(self._host_prof)).to(self.device, self.tdtype)
        self._keep += [cbd, pr_all]
        off = 0
        coff = 0
        poff = 0
        self._bufs = []
        for d in range(n):
            des = arr[d]
            ny, sy, m = des.ny, des.sy, sizes[d]

            def take():
                nonlocal off
                t = arena[off:off + m].view(ny, sy)
                off += m
                return t

            buf = {}
            for name in self._FIELD_ARRAYS:
                buf[name] = [take(), take()]
                getattr(des, name)[0] = buf[name][0].data_ptr()
                getattr(des, name)[1] = buf[name][1].data_ptr()
            for name in ("psi_ex", "psi_ey"):
                buf[name] = take()
                setattr(des, name, buf[name].data_ptr())
            des.cb = cbd[coff:coff + m].data_ptr()
            coff += m
            np_ = self._host_prof[d].size
            des.prof = pr_all[poff:poff + np_].data_ptr()
            poff += np_
            self._bufs.append(buf)
        self._host_cb = []
        self._host_prof = []

    def _build(self, nsteps: int):
        torch = _torch()
        n = len(self.sims)
        self._host_cb = []
        self._host_prof = []
        arr = (N.pm_batch_design * n)()
        times = []
        t = 0.0
        for _ in range(nsteps):
            t += self.dt
            times.append(t)
        times = np.array(times)
        col = 0
        self.mon = []
        tab_e = tab_h = None
        freqs = None
        for d, s in enumerate(self.sims):
            col, mons = self._design(s, arr[d], col)
            self.mon.append(mons)
            for m, _ in mons:
                f = np.asarray(m.frequencies, dtype=np.float64)
                if freqs is None:
                    freqs = f
                elif f.shape != freqs.shape or not np.array_equal(f, freqs):
                    raise NotImplementedError("every monitor of a batch shares one frequency list")
            e, h = self._sources(s, times)
            for got, name in ((e, "e"), (h, "h")):
                des_line = arr[d].src_e if name == "e" else arr[d].src_h
                if got is None:
                    des_line[0], des_line[1], des_line[2] = 0, 0, 0
                    continue
                des_line[0], des_line[1], des_line[2] = got[0], got[1], got[2]
                ref = tab_e if name == "e" else tab_h
                if ref is None:
                    if name == "e":
                        tab_e = got[3]
                    else:
                        tab_h = got[3]
                elif not np.array_equal(ref, got[3]):
                    raise NotImplementedError("every design of a batch shares its source waveform")
        self._place_arrays(arr)
        zeros = np.zeros(max(nsteps, 1))
        tab_e = zeros if tab_e is None else tab_e
        tab_h = zeros if tab_h is None else tab_h
        self.ncols = max(col, 1)
        self.freqs = freqs if freqs is not None else np.zeros(0)
        p = N.pm_batch_plan()
        p.abi_version = N.ABI_VERSION
        p.precision = self.precision
        p.coeff_kind = self.coeff_kind
        p.n_design = n
        p.dx, p.dy, p.dt = self.d[0], self.d[1], self.dt
        p.design = C.cast(arr, C.POINTER(N.pm_batch_design))
        te = torch.from_numpy(np.ascontiguousarray(tab_e)).to(self.device)
        th = torch.from_numpy(np.ascontiguousarray(tab_h)).to(self.device)
        self._keep += [te, th]
        p.src_e_table, p.src_h_table = te.data_ptr(), th.data_ptr()
        p.table_step0, p.table_len = 0, max(nsteps, 1)
        T = self.sample_T
        X = self._alloc((T, self.ncols))
        self._X = X
        p.X, p.ncols, p.sample_T = X.data_ptr(), self.ncols, T
        nf = len(self.freqs)
        p.nfreq = nf
        self.acc = None
        if nf:
            acc = [self._alloc((nf, self.ncols)), self._alloc((nf, self.ncols))]
            if self.precision == N.FP32:
                acc += [self._alloc((nf, self.ncols)), self._alloc((nf, self.ncols))]
            else:
                acc += [None, None]
            self.acc = acc
            g = p.dft
            g.col0, g.col1, g.f0, g.f1 = 0, self.ncols, 0, nf
            g.acc_re, g.acc_im = acc[0].data_ptr(), acc[1].data_ptr()
            g.cmp_re = acc[2].data_ptr() if acc[2] is not None else None
            g.cmp_im = acc[3].data_ptr() if acc[3] is not None else None
            w = np.exp(-1j * (2 * np.pi * self.freqs)[:, None] * times[None, :]) * self.dt
            Wre = torch.from_numpy(np.ascontiguousarray(w.real)).to(self.device, self.tdtype)
            Wim = torch.from_numpy(np.ascontiguousarray(w.imag)).to(self.device, self.tdtype)
            self._keep += [Wre, Wim]
            p.W_re, p.W_im = Wre.data_ptr(), Wim.data_ptr()
        p.chunk = self.chunk
        wsb = self.ctx.workspace_bytes(p)
        if self.auto_tune:
            from ..autotune import BATCH_CANDIDATES

            for c in BATCH_CANDIDATES["chunk"]:
                p.chunk = c
                wsb = max(wsb, self.ctx.workspace_bytes(p))
            p.chunk = self.chunk
        ws = self._alloc((wsb,), torch.uint8)
        p.work, p.work_bytes = ws.data_ptr(), wsb
        self.plan = p
        self._arr = arr
        # the library's stream runs after the fills and uploads queued on
        # torch's stream (a stream-side wait: other engines' work in flight
        # on the device is not waited for)
        self.stream.wait_stream(torch.cuda.current_stream(self.device))
        self.ctx.configure(p, keep=[arr])
        self.table_len = max(nsteps, 1)
        self._built = True

    def _zero_state(self):
        """Fields, CPML memory, sample rows and spectra back to zero (the
        state a configured run starts from), on the engine's stream."""
        torch = _torch()
        with torch.cuda.stream(self.stream):
            self._arena.zero_()
            self._X.zero_()
            for a in (self.acc or []):
                if a is not None:
                    a.zero_()

    def reconfigure(self, chunk: int):
        """Re-plan with ``chunk`` rows per tile; the step counter restarts."""
        self.chunk = int(chunk)
        self.plan.chunk = self.chunk
        self.stream.wait_stream(_torch().cuda.current_stream(self.device))
        self.ctx.configure(self.plan, keep=[self._arr])

    # ---------------------------------------------------------------- running
    def run_steps(self, nsteps: int, graph: bool = True, timing: bool = False) -> None:
        """Advance every design by ``nsteps`` steps (one configuration per
        engine: the tables cover exactly these steps).  ``timing`` records
        HIP-event times of every step launch / flush / graph replay."""
        self.start(nsteps, graph=graph, timing=timing)
        self.finish()

    def start(self, nsteps: int, graph: bool = True, timing: bool = False) -> None:
        """Set up and queue the whole run of ``nsteps`` steps on the engine's
        stream, returning without waiting for it; ``finish`` waits and
        delivers the results.  The host is free meanwhile (a sweep driver
        builds the next group of designs)."""
        if self._built:
            raise RuntimeError("a BatchEngine runs once; build a new one for another run")
        self._build(nsteps)
        if self.auto_tune:
            from ..autotune import tune_batch

            self.tuning = tune_batch(self)
        T = self.sample_T
        self.ctx.set_timing(timing)
        self.ctx.run(nsteps, graph=graph and T % 2 == 0)
        self.ctx.flush()
        self._queued = nsteps

    def finish(self) -> None:
        """Wait for the run ``start`` queued; set every Simulation's step
        count and time and fill its monitors."""
        nsteps = getattr(self, "_queued", None)
        if nsteps is None:
            raise RuntimeError("finish() without start()")
        self._queued = None
        self.ctx.sync()
        # per design: did its fields or spectra leave the finite numbers?
        st = self.ctx.design_status(len(self.sims))
        self.diverged = [bool(v & 1) for v in st]
        self.step_count = nsteps
        t = 0.0
        for _ in range(nsteps):
            t += self.dt
        for s in self.sims:
            s.step_count = nsteps
            s.current_time = t
        self._ingest()

    def _ingest(self):
        if self.acc is None:
            return
        re, im, cre, cim = self.acc
        if cre is not None:
            R = (re.double() + cre.double()).cpu().numpy()
            I = (im.double() + cim.double()).cpu().numpy()
        else:
            R, I = re.cpu().numpy(), im.cpu().numpy()
        for d, mons in enumerate(self.mon):
            for m, cols in mons:
                if not all(np.isfinite(R[:, c0:c1]).all() and np.isfinite(I[:, c0:c1]).all()
                           for c0, c1 in cols.values()):
                    self.diverged[d] = True
                m._ingest_spectra({c: R[:, c0:c1] + 1j * I[:, c0:c1] for c, (c0, c1) in cols.items()})

    def fields(self, d: int) -> dict:
        """Current Ez, Hx, Hy of design d in the reference's [i, j] layout."""
        nx, ny = self._arr[d].nx, self._arr[d].ny
        q = self.ctx.parity()
        b = self._bufs[d]
        ext = {"ez": (nx - 1, ny - 1), "hx": (nx - 1, ny), "hy": (nx, ny - 1)}
        out = {}
        for name, comp in (("ez", "Ez"), ("hx", "Hx"), ("hy", "Hy")):
            a = b[name][q].cpu().numpy()
            ex, ey = ext[name]
            out[comp] = a[:ey, :ex].T.astype(np.float64)
        return out

    def timing(self, which: int):
        return self.ctx.timing(which)
