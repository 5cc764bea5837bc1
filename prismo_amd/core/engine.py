"""Binds a field set to a libpmfdtd context: coefficients, CPML, ADE state,
source tables, sample/DFT buffers, all as torch device tensors whose
pointers go into one ``pm_plan``.

This is the native side of ``MaxwellUpdater.step`` / ``FDTDSolver.run_steps``
/ ``Simulation.run`` (src/prismo/core/solver.py:535-552, 643-662;
src/prismo/core/simulation.py:107-164) on the 'mi355x' backend.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .. import _native as N
from .coefficients import cell_coefficients, component_average, is_lossless_nonmagnetic
from .grid import COMPONENTS

EPS0 = 8.854187817e-12


@dataclass
class AdeRegion:
    """One dispersive region on one E component (index box, coefficients)."""

    comp: str
    box: tuple
    kind: int            # 1 Lorentz, 2 Drude
    eps_inf: float
    c: tuple             # (c0, c1, c2, c3)


@dataclass
class _MonitorBinding:
    monitor: object
    spec: dict
    cols: dict = field(default_factory=dict)  # comp -> (col0, col1)
    f0: int = 0
    acc: Optional[tuple] = None               # (re, im, cre, cim) device tensors


def _torch():
    import torch

    return torch


class StepEngine:
    def __init__(self, grid, dt, fields, coeffs, *, scheme: str = "compat", cpml=None,
                 ade=(), coefficient_storage: str = "auto", fused: bool = True):
        torch = _torch()
        b = fields.backend
        if getattr(b, "name", None) != "mi355x":
            raise NotImplementedError(
                "prismo_amd runs the time step only on the 'mi355x' backend "
                f"(fields live on '{getattr(b, 'name', b)}'); call set_backend('mi355x')"
            )
        self.grid = grid
        self.dt = float(dt)
        self.fields = fields
        self.backend = b
        self.device = b.device
        self.is_2d = grid.is_2d
        fdt = fields["Ez"].dtype
        if fdt == torch.float64:
            self.precision = N.FP64
        elif fdt == torch.float32:
            self.precision = N.FP32
        else:
            raise TypeError(f"fields must be float32 or float64, got {fdt}")
        self.tdtype = fdt
        if scheme not in ("compat", "yee"):
            raise ValueError("scheme must be 'compat' or 'yee'")
        self.scheme = scheme
        if scheme == "compat" and self.precision != N.FP64:
            raise NotImplementedError("the compat scheme reproduces the FP64 reference only")
        if scheme == "compat" and (cpml is not None or ade):
            raise NotImplementedError("CPML and ADE exist only in the 'yee' scheme")
        self._keep: list = []
        self._bindings: dict = {}
        self.plan = N.pm_plan()
        self._fill_grid()
        self._make_arena(cpml)
        self._fill_coefficients(coeffs, coefficient_storage)
        self._fill_cpml(cpml)
        self._fill_ade(list(ade))
        self._fill_fused(fused)
        self._scratch_ptr = self._alloc((N.SCRATCH_BYTES // 4,), torch.int32).data_ptr()
        self.plan.scratch = self._scratch_ptr
        self.static_scratch = False
        stream = torch.cuda.current_stream(self.device)
        self.ctx = N.Context(self.device.index or 0, stream.cuda_stream)
        self._io_keep: list = []
        self._src_arr = None
        self._sbox_arr = None
        self._dft_arr = None
        self.configure_io()

    # ------------------------------------------------------------------ setup
    def _alloc(self, shape, dtype):
        t = _torch().zeros(shape, dtype=dtype, device=self.device)
        self._keep.append(t)
        return t

    def _make_arena(self, cpml):
        """One zeroed allocation for the coefficient, CPML memory and profile
        arrays, so they sit within a few hundred MB of each other: the fused
        kernel addresses them with 32-bit offsets from one base."""
        nx, ny, nz = self.grid.dimensions
        esz = 8 if self.precision == N.FP64 else 4
        field = nz * ny * self.sy * esz
        total = 3 * field + 6 * (nx + ny + nz) * esz + 4096
        if cpml is not None:
            cells = nx * ny * nz
            for a, n in enumerate((nx, ny, nz)):
                S = cpml.thickness[2 * a] + cpml.thickness[2 * a + 1]
                total += 8 * (S * cells // n) * esz + 8 * 256
        self._arena = _torch().zeros((total,), dtype=_torch().uint8, device=self.device)
        self._arena_off = 0

    def _carve(self, shape, dtype):
        """Zeroed array from the arena (256-byte aligned), else a plain one."""
        torch = _torch()
        n = int(np.prod(shape))
        nbytes = n * torch.empty((), dtype=dtype).element_size()
        start = (self._arena_off + 255) // 256 * 256
        if start + nbytes > self._arena.numel():
            return self._alloc(shape, dtype)
        t = self._arena[start:start + nbytes].view(dtype).view(shape)
        self._arena_off = start + nbytes
        self._keep.append(t)
        return t

    def _padded(self, host: np.ndarray, comp: str, dtype, buf=None):
        """Device array in the shared padded layout holding ``host`` (shaped
        like the component) at its reference indices, zero elsewhere (into
        ``buf`` when given)."""
        torch = _torch()
        nx, ny, nz = self.grid.dimensions
        if buf is None:
            buf = self._alloc((nz, ny, self.sy), dtype)
        if self.is_2d:
            view = buf[0, : host.shape[1], : host.shape[0]].permute(1, 0)
        else:
            view = buf[: host.shape[2], : host.shape[1], : host.shape[0]].permute(2, 1, 0)
        view.copy_(torch.from_numpy(np.ascontiguousarray(host)).to(self.device, dtype))
        return buf

    def _fill_grid(self):
        p = self.plan
        g = self.grid
        nx, ny, nz = g.dimensions
        ez = self.fields["Ez"]
        st = ez.stride()
        self.sy = st[1]
        self.sz = self.sy * ny
        for c in COMPONENTS:
            f = self.fields[c]
            s = f.stride()
            if s[0] != 1 or s[1] != self.sy or (not self.is_2d and s[2] != self.sz):
                raise ValueError(
                    f"field {c} is not in the padded x-fastest layout "
                    "(allocate fields through ElectromagneticFields on the mi355x backend)"
                )
            if f.device != self.device:
                raise ValueError(f"field {c} is on {f.device}, expected {self.device}")
        p.abi_version = N.ABI_VERSION
        p.ndim = 2 if self.is_2d else 3
        p.scheme = N.SCHEME_COMPAT if self.scheme == "compat" else N.SCHEME_YEE
        p.precision = self.precision
        p.nx, p.ny, p.nz = nx, ny, nz
        p.sy, p.sz = self.sy, self.sz
        p.dx, p.dy = g.dx, g.dy
        p.dz = g.dz if not self.is_2d else 0.0
        p.dt = self.dt
        for a, c in enumerate(("Ex", "Ey", "Ez")):
            p.E[a] = self.fields[c].data_ptr()
        for a, c in enumerate(("Hx", "Hy", "Hz")):
            p.H[a] = self.fields[c].data_ptr()
        if self.scheme == "yee":
            self.need_e = ["Ex", "Ey", "Ez"] if not self.is_2d else ["Ez"]
            self.need_h = ["Hx", "Hy", "Hz"] if not self.is_2d else ["Hx", "Hy"]
        else:
            self.need_e = ["Ex", "Ey", "Ez"]
            self.need_h = ["Hx", "Hy", "Hz"]

    def _fill_coefficients(self, coeffs, storage):
        if len(coeffs) == 3 and isinstance(coeffs[0], str) and coeffs[0] == "device":
            return self._fill_coefficients_device(coeffs[1], coeffs[2], storage)
        torch = _torch()
        Ca, Cb, Da, Db, mu_rel, sig_e, sig_m = coeffs
        p = self.plan
        lossless = is_lossless_nonmagnetic(mu_rel, sig_e, sig_m)
        if storage == "fp16":
            if not (lossless and self.precision == N.FP32 and self.scheme == "yee"):
                raise ValueError("FP16 coefficients need lossless media, FP32 fields, 'yee'")
            kind = N.COEFF_CB_F16
        elif storage in ("auto", "native"):
            kind = N.COEFF_CB if lossless else N.COEFF_FULL
        elif storage == "full":
            kind = N.COEFF_FULL
        else:
            raise ValueError("coefficient_storage must be 'auto', 'full' or 'fp16'")
        self.coeff_kind = kind
        p.coeff_kind = kind
        sl = (lambda A: A[:, :, 0]) if self.is_2d else (lambda A: A)
        self.cb_scale = self.dt / EPS0
        p.cb_scale = self.cb_scale
        if kind != N.COEFF_FULL:
            d0 = float(np.asarray(Db).flat[0])
            if not np.all(Db == d0):
                raise ValueError("uniform Db expected for lossless non-magnetic media")
            p.db_const = d0
        self.cb_host = {}
        self.cb_dev = {}
        nx, ny, nz = self.grid.dimensions
        cb_dt = torch.float16 if kind == N.COEFF_CB_F16 else self.tdtype
        # the three Cb arrays back to back (the fused kernel's layout)
        cb_all = None if self.is_2d else self._carve((3, nz, ny, self.sy), cb_dt)
        for a, c in enumerate(("Ex", "Ey", "Ez")):
            if c not in self.need_e:
                continue
            cb = component_average(sl(Cb), c, self.is_2d)
            dst = None if cb_all is None else cb_all[a]
            if kind == N.COEFF_CB_F16:
                q = (cb / self.cb_scale).astype(np.float16)
                self.cb_host[c] = q.astype(np.float64) * self.cb_scale
                self.cb_dev[c] = self._padded(q, c, torch.float16, dst)
            else:
                self.cb_host[c] = cb
                self.cb_dev[c] = self._padded(cb, c, self.tdtype, dst)
            p.cb[a] = self.cb_dev[c].data_ptr()
            if kind == N.COEFF_FULL:
                p.ca[a] = self._padded(component_average(sl(Ca), c, self.is_2d), c,
                                       self.tdtype).data_ptr()
        if kind == N.COEFF_FULL:
            for a, c in enumerate(("Hx", "Hy", "Hz")):
                if c not in self.need_h:
                    continue
                p.da[a] = self._padded(component_average(sl(Da), c, self.is_2d), c,
                                       self.tdtype).data_ptr()
                p.db[a] = self._padded(component_average(sl(Db), c, self.is_2d), c,
                                       self.tdtype).data_ptr()

    def _fill_coefficients_device(self, shapes, background_eps, storage):
        """Cb from shapes painted on the device (include/pmgeom.h): a
        material index per node, then per E component the average of the
        materials' node Cb, straight into the padded arrays.  Same bits as
        the host path on the equivalent node arrays."""
        from ..geometry import raster
        from ..geometry.shapes import Material

        torch = _torch()
        if self.scheme != "yee":
            raise NotImplementedError("device-painted shapes need scheme='yee'")
        if storage == "fp16":
            if self.precision != N.FP32:
                raise ValueError("FP16 coefficients need FP32 fields")
            kind = N.COEFF_CB_F16
        elif storage in ("auto", "native"):
            kind = N.COEFF_CB
        else:
            raise ValueError("device-painted shapes take coefficient_storage 'auto' or 'fp16'")
        p = self.plan
        self.coeff_kind = kind
        p.coeff_kind = kind
        self.cb_scale = self.dt / EPS0
        p.cb_scale = self.cb_scale
        mats = raster.material_table(shapes, Material("background", background_eps))
        eps = np.array([m.epsilon_r for m in mats], dtype=np.float64)
        one, zero = np.ones_like(eps), np.zeros_like(eps)
        _, cb_lut, _, db = cell_coefficients(eps, one, zero, zero, self.dt)
        p.db_const = float(db[0])
        g = self.grid
        nx, ny, nz = g.dimensions
        x = g.origin[0] + np.arange(nx) * g.dx
        y = g.origin[1] + np.arange(ny) * g.dy
        z = np.zeros(1) if self.is_2d else g.origin[2] + np.arange(nz) * g.dz
        idx, _ = raster.paint(shapes, x, y, z, self.device, mats)
        self.material_index = idx
        self.materials = mats
        self.cb_host = {}
        self.cb_dev = {}
        cb_dt = torch.float16 if kind == N.COEFF_CB_F16 else self.tdtype
        store = {N.COEFF_CB_F16: N.CB_F16}.get(kind, N.CB_F64 if self.precision == N.FP64
                                              else N.CB_F32)
        cb_all = None if self.is_2d else self._carve((3, nz, ny, self.sy), cb_dt)
        for a, c in enumerate(("Ex", "Ey", "Ez")):
            if c not in self.need_e:
                continue
            buf = self._alloc((nz, ny, self.sy), cb_dt) if cb_all is None else cb_all[a]
            raster.cb_from_index(idx, cb_lut, a, store, self.cb_scale, buf, self.sy, self.sz)
            self.cb_dev[c] = buf
            p.cb[a] = buf.data_ptr()

    def coefficient_at(self, comp: str, box) -> tuple:
        """(Cb, Db) in FP64 at the cells of ``box`` ((i0, i1), (j0, j1),
        (k0, k1)) of E component ``comp`` and of the matching H points, read
        back from the device arrays (lossless media: Db is one constant)."""
        if self.coeff_kind == N.COEFF_FULL:
            raise NotImplementedError("coefficients of lossy or magnetic media")
        (i0, i1), (j0, j1), (k0, k1) = box
        t = self.cb_dev[comp][k0:k1, j0:j1, i0:i1].permute(2, 1, 0).double().cpu().numpy()
        if self.coeff_kind == N.COEFF_CB_F16:
            t = t * self.cb_scale
        if self.is_2d:
            t = t[:, :, 0]
        return t, float(self.plan.db_const)

    def _fill_cpml(self, cpml):
        p = self.plan
        self.cpml = cpml
        if cpml is None:
            return
        prof = cpml.profiles(self.dt)
        nx, ny, nz = self.grid.dimensions
        n = (nx, ny, nz)
        for a in range(3):
            tl, th = cpml.thickness[2 * a], cpml.thickness[2 * a + 1]
            p.pml[2 * a], p.pml[2 * a + 1] = tl, th
            if tl + th == 0:
                continue
            pr = prof[a]
            t = _torch()
            # the six profiles of an axis back to back, then the two CPML
            # memory arrays of each side back to back (the fused kernel's layout)
            keys = (("be", "pml_be"), ("ce", "pml_ce"), ("ike", "pml_ike"),
                    ("bh", "pml_bh"), ("ch", "pml_ch"), ("ikh", "pml_ikh"))
            pbuf = self._carve((6, n[a]), self.tdtype)
            pbuf.copy_(t.from_numpy(np.stack([pr[k] for k, _ in keys])).to(self.device, self.tdtype))
            for r, (_, attr) in enumerate(keys):
                getattr(p, attr)[a] = pbuf[r].data_ptr()
            S = tl + th
            shape = {0: (nz, ny, S), 1: (nz, S, nx), 2: (S, ny, nx)}[a]
            for side, need, comps in ((p.psi_e, self.need_e, ("Ex", "Ey", "Ez")),
                                      (p.psi_h, self.need_h, ("Hx", "Hy", "Hz"))):
                want = [comps[(a + 2) % 3 if q == 1 else (a + 1) % 3] in need for q in range(2)]
                if all(want):
                    pair = self._carve((2,) + shape, self.tdtype)
                    for q in range(2):
                        self._keep.append(pair[q])
                        side[a][q] = pair[q].data_ptr()
                else:
                    for q in range(2):
                        if want[q]:
                            side[a][q] = self._alloc(shape, self.tdtype).data_ptr()

    def _fill_fused(self, fused: bool):
        """Second buffer set for the fused ping-pong step (3-D yee, one x-row
        per wave64: pitch / 64 in {2, 4, 8}; 2-D TM yee: Ez, Hx, Hy)."""
        p = self.plan
        self.fused = False
        if not fused or self.scheme != "yee":
            return
        if self.is_2d:
            self._fill_fused_2d()
            return
        if self.sy not in (128, 256, 512):
            return
        nx, ny, nz = self.grid.dimensions
        # second field set: Ex Ey Ez Hx Hy Hz back to back at the stride of
        # the first set (ElectromagneticFields allocates it the same way)
        setb = self._alloc((6, nz, ny, self.sy), self.tdtype)
        for a in range(3):
            p.E2[a] = setb[a].data_ptr()
            p.H2[a] = setb[3 + a].data_ptr()
        for a in range(3):
            S = p.pml[2 * a] + p.pml[2 * a + 1]
            shape = {0: (nz, ny, S), 1: (nz, S, nx), 2: (S, ny, nx)}[a]
            for src, dst in ((p.psi_h, p.psi_h2), (p.psi_e, p.psi_e2)):
                if src[a][0] and src[a][1]:
                    pair = self._carve((2,) + shape, self.tdtype)
                    for q in range(2):
                        self._keep.append(pair[q])
                        dst[a][q] = pair[q].data_ptr()
                else:
                    for q in range(2):
                        if src[a][q]:
                            dst[a][q] = self._alloc(shape, self.tdtype).data_ptr()
        self.fused = True

    def _fill_fused_2d(self):
        """2-D TM (the yee scheme steps Ez, Hx, Hy in 2-D): a second Ez, Hx,
        Hy and second H-side CPML memory (Hy in the x slabs, Hx in the y
        slabs); lossy media and dispersive boxes keep the unfused kernels."""
        p = self.plan
        if p.n_ade or p.coeff_kind == N.COEFF_FULL:
            return
        nx, ny, _ = self.grid.dimensions
        setb = self._alloc((3, 1, ny, self.sy), self.tdtype)
        p.E2[2] = setb[0].data_ptr()
        p.H2[0] = setb[1].data_ptr()
        p.H2[1] = setb[2].data_ptr()
        for a in range(2):
            S = p.pml[2 * a] + p.pml[2 * a + 1]
            shape = {0: (1, ny, S), 1: (1, S, nx)}[a]
            for q in range(2):
                if p.psi_h[a][q]:
                    p.psi_h2[a][q] = self._alloc(shape, self.tdtype).data_ptr()
        self.fused = True

    def _fill_ade(self, regions):
        p = self.plan
        if len(regions) > 8:
            raise ValueError("at most 8 dispersive boxes are supported")
        arr = (N.pm_ade_box * max(1, len(regions)))()
        self.ade_state = []
        for i, r in enumerate(regions):
            (i0, i1), (j0, j1), (k0, k1) = r.box
            cells = max(0, i1 - i0) * max(0, j1 - j0) * max(0, k1 - k0)
            P = self._alloc((max(cells, 1),), self.tdtype)
            Q = self._alloc((max(cells, 1),), self.tdtype) if r.kind == 1 else None
            self.ade_state.append((P, Q))
            a = arr[i]
            a.comp = ("Ex", "Ey", "Ez").index(r.comp)
            a.kind = r.kind
            a.i0, a.i1, a.j0, a.j1, a.k0, a.k1 = i0, i1, j0, j1, k0, k1
            a.eps_inf = r.eps_inf
            a.c0, a.c1, a.c2, a.c3 = r.c
            a.cb_scale = self.dt / EPS0
            a.p = P.data_ptr()
            a.p_prev = Q.data_ptr() if Q is not None else None
        self._ade_arr = arr
        p.n_ade = len(regions)
        p.ade = C.cast(arr, C.POINTER(N.pm_ade_box))

    # ---------------------------------------------------------------- I/O
    def configure_io(self, terms=(), monitors=(), times=None, step0: int = 0,
                     sample_T: int = 128):
        """(Re)configure the plan with source terms, monitors and the step
        tables for steps [step0, step0 + len(times))."""
        torch = _torch()
        p = self.plan
        self._io_keep = []
        terms = [t for t in terms if _box_cells(t.box) > 0]
        if len(terms) > 16:
            raise ValueError("at most 16 device source terms per run")
        src = (N.pm_source * max(1, len(terms)))()
        for i, t in enumerate(terms):
            s = src[i]
            s.comp = N.COMP_ID[t.comp]
            s.kind = 1 if t.profile is not None else 0
            (s.i0, s.i1), (s.j0, s.j1), (s.k0, s.k1) = t.box
            tab = torch.from_numpy(np.ascontiguousarray(t.table, dtype=np.float64)).to(self.device)
            self._io_keep.append(tab)
            s.table = tab.data_ptr()
            if t.profile is not None:
                pr = torch.from_numpy(np.ascontiguousarray(t.profile, dtype=np.float64)).to(self.device)
                self._io_keep.append(pr)
                s.profile = pr.data_ptr()
            s.divisor = float(t.divisor)
        self._src_arr = src
        p.n_src = len(terms)
        p.src = C.cast(src, C.POINTER(N.pm_source))
        nsteps = 0 if times is None else len(times)
        p.table_step0 = step0
        p.table_len = max(nsteps, 1)

        # monitors -> sample boxes, DFT groups
        boxes = []
        groups = []
        freqs_all = []
        col = 0
        self.mon = []
        for m in monitors:
            key = id(m)
            spec = m.device_spec()
            bnd = self._bindings.get(key)
            if bnd is None or bnd.monitor is not m:
                bnd = _MonitorBinding(m, spec)
                self._bindings[key] = bnd
            bnd.spec = spec
            c0 = col
            for comp, box, _shape in spec["boxes"]:
                ncell = _box_cells(box)
                if ncell == 0:
                    bnd.cols[comp] = (col, col)
                    continue
                if comp not in self.need_e + self.need_h:
                    raise NotImplementedError(f"{comp} is not simulated by this scheme")
                boxes.append((N.COMP_ID[comp], box, col))
                bnd.cols[comp] = (col, col + ncell)
                col += ncell
            if spec.get("freqs") is not None and len(spec["freqs"]) and col > c0:
                nf = len(spec["freqs"])
                bnd.f0 = len(freqs_all)
                freqs_all.extend(spec["freqs"].tolist())
                if bnd.acc is None or bnd.acc[0].shape != (nf, col - c0):
                    acc_dt = torch.float64 if self.precision == N.FP64 else torch.float32
                    re = self._alloc((nf, col - c0), acc_dt)
                    im = self._alloc((nf, col - c0), acc_dt)
                    cre = cim = None
                    if self.precision == N.FP32:
                        cre = self._alloc((nf, col - c0), acc_dt)
                        cim = self._alloc((nf, col - c0), acc_dt)
                    bnd.acc = (re, im, cre, cim)
                groups.append((c0, col, bnd.f0, bnd.f0 + nf, bnd.acc))
            self.mon.append(bnd)
        if len(boxes) > 24:
            raise ValueError("at most 24 monitored component boxes per run")
        sb = (N.pm_sample_box * max(1, len(boxes)))()
        for i, (comp, box, c0) in enumerate(boxes):
            (sb[i].i0, sb[i].i1), (sb[i].j0, sb[i].j1), (sb[i].k0, sb[i].k1) = box
            sb[i].comp = comp
            sb[i].col0 = c0
        self._sbox_arr = sb
        p.n_sbox = len(boxes)
        p.sbox = C.cast(sb, C.POINTER(N.pm_sample_box))
        p.ncols = col
        self.ncols = col
        self.sample_T = int(sample_T)
        p.sample_T = self.sample_T
        if boxes:
            xdt = torch.float64 if self.precision == N.FP64 else torch.float32
            X = torch.zeros((self.sample_T, col), dtype=xdt, device=self.device)
            self._io_keep.append(X)
            self.X = X
            p.X = X.data_ptr()
        else:
            self.X = None
            p.X = None
        dg = (N.pm_dft_group * max(1, len(groups)))()
        for i, (c0, c1, f0, f1, acc) in enumerate(groups):
            d = dg[i]
            d.col0, d.col1, d.f0, d.f1 = c0, c1, f0, f1
            d.acc_re, d.acc_im = acc[0].data_ptr(), acc[1].data_ptr()
            d.cmp_re = acc[2].data_ptr() if acc[2] is not None else None
            d.cmp_im = acc[3].data_ptr() if acc[3] is not None else None
        self._dft_arr = dg
        p.n_dft = len(groups)
        p.dft = C.cast(dg, C.POINTER(N.pm_dft_group))
        p.nfreq_total = len(freqs_all)
        if groups:
            w = np.exp(-1j * (2 * np.pi * np.asarray(freqs_all))[:, None] * np.asarray(times)[None, :]) * self.dt
            wdt = torch.float64 if self.precision == N.FP64 else torch.float32
            Wre = torch.from_numpy(np.ascontiguousarray(w.real)).to(self.device, wdt)
            Wim = torch.from_numpy(np.ascontiguousarray(w.imag)).to(self.device, wdt)
            self._io_keep += [Wre, Wim]
            p.W_re, p.W_im = Wre.data_ptr(), Wim.data_ptr()
        else:
            p.W_re = p.W_im = None
        if self.static_scratch:
            p.scratch = None
            self.ctx.configure_static(self.plan)
        else:
            p.scratch = self._scratch_ptr
            self.ctx.configure(self.plan)

    def use_static_scratch(self, on: bool = True):
        """Configure through pm_configure_static (the library's per-device
        scratch; no device allocation in the configure call) from now on.
        Every engine doing so on one device shares that scratch: configure
        and run such engines one after another, never interleaved."""
        self.static_scratch = bool(on)
        self.configure_io()

    # ---------------------------------------------------------------- stepping
    def update_h(self):
        self.ctx.update_h()

    def update_e(self):
        self.ctx.update_e()

    def run(self, step0: int, nsteps: int, flags: int = 0):
        if nsteps > 0:
            self.ctx.run(step0, nsteps, flags)

    def flush(self):
        self.ctx.dft_flush()

    def sync(self):
        self.ctx.sync()

    def sample_rows(self, row0: int, nrows: int) -> np.ndarray:
        return self.X[row0:row0 + nrows].double().cpu().numpy()

    def spectra(self, bnd: _MonitorBinding) -> dict:
        """Accumulated spectra of one monitor as {comp: complex [nf, ncells]}."""
        re, im, cre, cim = bnd.acc
        if cre is not None:
            R = re.double() + cre.double()
            I = im.double() + cim.double()
        else:
            R, I = re, im
        R = R.cpu().numpy()
        I = I.cpu().numpy()
        c_base = min(c0 for c0, _ in bnd.cols.values())
        out = {}
        for comp, (c0, c1) in bnd.cols.items():
            out[comp] = R[:, c0 - c_base:c1 - c_base] + 1j * I[:, c0 - c_base:c1 - c_base]
        return out

    def reset_state(self):
        """Zero CPML and ADE memory and spectral accumulators (fields are the
        caller's)."""
        for t in self._psi_tensors():
            t.zero_()
        for P, Q in getattr(self, "ade_state", []):
            P.zero_()
            if Q is not None:
                Q.zero_()
        for b in self._bindings.values():
            if b.acc is not None:
                for t in b.acc:
                    if t is not None:
                        t.zero_()

    def _psi_tensors(self):
        ptrs = set()
        p = self.plan
        for a in range(3):
            for q in range(2):
                for v in (p.psi_e[a][q], p.psi_h[a][q], p.psi_h2[a][q], p.psi_e2[a][q]):
                    if v:
                        ptrs.add(v)
        return [t for t in self._keep if t.data_ptr() in ptrs]


def _box_cells(box) -> int:
    (i0, i1), (j0, j1), (k0, k1) = box
    return max(0, i1 - i0) * max(0, j1 - j0) * max(0, k1 - k0)
