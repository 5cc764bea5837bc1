"""ctypes binding of libpmfdtd (include/pmfdtd.h).

The structures below mirror the header field for field; ``check_layout``
compares their sizes with what the library was built with through
``pm_abi_version``.  Loading fails loudly: there is no CPU fallback for the
time step.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libpmfdtd.so"
ABI_VERSION = 7
SCRATCH_BYTES = 4096

PM_EX, PM_EY, PM_EZ, PM_HX, PM_HY, PM_HZ = range(6)
COMP_ID = {"Ex": 0, "Ey": 1, "Ez": 2, "Hx": 3, "Hy": 4, "Hz": 5}
SCHEME_COMPAT, SCHEME_YEE = 0, 1
FP64, FP32 = 0, 1
COEFF_FULL, COEFF_CB, COEFF_CB_F16 = 0, 1, 2
RUN_SOURCES, RUN_SAMPLE, RUN_DFT, RUN_GRAPH, RUN_IO_ONLY = 1, 2, 4, 8, 16

ERRORS = {
    -1: "invalid argument",
    -2: "not supported",
    -3: "no plan configured",
    -4: "step outside uploaded tables",
}

vp = C.c_void_p
i32 = C.c_int32
i64 = C.c_int64
f64 = C.c_double


class pm_source(C.Structure):
    _fields_ = [
        ("comp", i32), ("kind", i32),
        ("i0", i32), ("i1", i32), ("j0", i32), ("j1", i32), ("k0", i32), ("k1", i32),
        ("table", vp), ("profile", vp), ("divisor", f64),
    ]


class pm_sample_box(C.Structure):
    _fields_ = [
        ("comp", i32),
        ("i0", i32), ("i1", i32), ("j0", i32), ("j1", i32), ("k0", i32), ("k1", i32),
        ("col0", i32),
    ]


class pm_dft_group(C.Structure):
    _fields_ = [
        ("col0", i32), ("col1", i32), ("f0", i32), ("f1", i32),
        ("acc_re", vp), ("acc_im", vp), ("cmp_re", vp), ("cmp_im", vp),
    ]


class pm_ade_box(C.Structure):
    _fields_ = [
        ("comp", i32), ("kind", i32),
        ("i0", i32), ("i1", i32), ("j0", i32), ("j1", i32), ("k0", i32), ("k1", i32),
        ("eps_inf", f64), ("c0", f64), ("c1", f64), ("c2", f64), ("c3", f64),
        ("cb_scale", f64), ("p", vp), ("p_prev", vp),
    ]


class pm_plan(C.Structure):
    _fields_ = [
        ("abi_version", i32), ("ndim", i32),
        ("scheme", i32), ("precision", i32), ("coeff_kind", i32), ("pad0", i32),
        ("nx", i32), ("ny", i32), ("nz", i32), ("pad1", i32),
        ("sy", i64), ("sz", i64),
        ("dx", f64), ("dy", f64), ("dz", f64), ("dt", f64),
        ("E", vp * 3), ("H", vp * 3),
        ("ca", vp * 3), ("cb", vp * 3), ("da", vp * 3), ("db", vp * 3),
        ("cb_scale", f64), ("db_const", f64),
        ("pml", i32 * 6),
        ("pml_be", vp * 3), ("pml_ce", vp * 3), ("pml_ike", vp * 3),
        ("pml_bh", vp * 3), ("pml_ch", vp * 3), ("pml_ikh", vp * 3),
        ("psi_e", (vp * 2) * 3), ("psi_h", (vp * 2) * 3),
        ("n_ade", i32), ("pad2", i32), ("ade", C.POINTER(pm_ade_box)),
        ("n_src", i32), ("pad3", i32), ("src", C.POINTER(pm_source)),
        ("table_step0", i64), ("table_len", i64),
        ("n_sbox", i32), ("sample_T", i32), ("sbox", C.POINTER(pm_sample_box)),
        ("ncols", i32), ("n_dft", i32), ("X", vp), ("dft", C.POINTER(pm_dft_group)),
        ("W_re", vp), ("W_im", vp),
        ("nfreq_total", i32), ("pad4", i32),
        ("scratch", vp),
        ("E2", vp * 3), ("H2", vp * 3), ("psi_h2", (vp * 2) * 3),
        ("psi_e2", (vp * 2) * 3),
    ]


# include/pmbatch.h
BATCH_MAX_LINES = 8


class pm_batch_line(C.Structure):
    _fields_ = [("comp", i32), ("i", i32), ("j0", i32), ("j1", i32), ("col", i32), ("pad", i32)]


class pm_batch_design(C.Structure):
    _fields_ = [
        ("nx", i32), ("ny", i32), ("sy", i32), ("n_lines", i32),
        ("pml", i32 * 4),
        ("ez", vp * 2), ("hx", vp * 2), ("hy", vp * 2),
        ("psi_hx", vp * 2), ("psi_hy", vp * 2), ("psi_ex", vp), ("psi_ey", vp),
        ("cb", vp), ("cb_scale", f64), ("db", f64), ("prof", vp),
        ("src_e", i32 * 3), ("src_h", i32 * 3),
        ("line", pm_batch_line * BATCH_MAX_LINES),
    ]


class pm_batch_plan(C.Structure):
    _fields_ = [
        ("abi_version", i32), ("precision", i32), ("coeff_kind", i32), ("n_design", i32),
        ("dx", f64), ("dy", f64), ("dt", f64),
        ("design", C.POINTER(pm_batch_design)),
        ("src_e_table", vp), ("src_h_table", vp),
        ("table_step0", i64), ("table_len", i64),
        ("X", vp), ("ncols", i32), ("sample_T", i32),
        ("dft", pm_dft_group),
        ("W_re", vp), ("W_im", vp),
        ("nfreq", i32), ("chunk", i32),
        ("work", vp), ("work_bytes", i64),
    ]


class pm_shape(C.Structure):
    _fields_ = [
        ("kind", i32), ("material", i32), ("axis", i32), ("n_vert", i32),
        ("vert0", i64),
        ("center", f64 * 3), ("half", f64 * 3),
        ("radius", f64), ("half_height", f64), ("zmin", f64), ("zmax", f64),
    ]


CB_F64, CB_F32, CB_F16 = 0, 1, 2
GEOM_SIGNATURES = {
    "pm_raster": ([vp, i32, i32, i32, vp, vp, vp, vp, i32, vp, i64, vp], C.c_int),
    "pm_cb_from_index": ([vp, i32, i32, i32, vp, vp, i32, i32, i32, f64, i64, i64, vp, vp],
                         C.c_int),
}

# Exported symbols and their signatures (the full set declared in the headers).
MLP_LINEAR, MLP_GELU, MLP_DGELU, MLP_PLAIN = 0, 1, 2, 3
fp = C.POINTER(C.c_float)
MLP_SIGNATURES = {
    "pm_mlp_gemm": ([vp, C.c_int, i32, i32, i32, vp, i64, i64, vp, i64, i64, vp, i64, vp, vp],
                    C.c_int),
    "pm_mlp_colsum": ([vp, i32, i32, vp, i64, vp], C.c_int),
    "pm_mlp_mse": ([vp, i32, i32, vp, vp, vp, vp, vp], C.c_int),
    "pm_mlp_adam": ([vp, i64, vp, vp, vp, vp, C.c_float, C.c_float, C.c_float, C.c_float,
                     C.c_float, i32], C.c_int),
    "pm_mlp_gemm_batched": ([vp, C.c_int, i32, i32, i32, i32, vp, i64, i64, i64, vp, i64, i64,
                             i64, vp, i64, i64, vp, i64, vp, i64], C.c_int),
    "pm_mlp_colsum_batched": ([vp, i32, i32, i32, vp, i64, i64, vp, i64], C.c_int),
    "pm_mlp_mse_batched": ([vp, i32, i32, i32, vp, vp, i64, vp, vp, vp], C.c_int),
}

BATCH_SIGNATURES = {
    "pm_batch_create": ([C.c_int, vp, C.POINTER(vp)], C.c_int),
    "pm_batch_destroy": ([vp], C.c_int),
    "pm_batch_last_error": ([vp], C.c_char_p),
    "pm_batch_workspace_bytes": ([C.POINTER(pm_batch_plan), C.POINTER(i64)], C.c_int),
    "pm_batch_configure": ([vp, C.POINTER(pm_batch_plan)], C.c_int),
    "pm_batch_run": ([vp, i32, C.c_uint32], C.c_int),
    "pm_batch_flush": ([vp], C.c_int),
    "pm_batch_parity": ([vp], C.c_int),
    "pm_batch_steps": ([vp], i64),
    "pm_batch_sync": ([vp], C.c_int),
    "pm_batch_design_status": ([vp, C.POINTER(i32), C.c_int32], C.c_int),
    "pm_batch_set_tuning": ([vp, C.c_int, C.c_int], C.c_int),
    "pm_batch_set_timing": ([vp, C.c_int], C.c_int),
    "pm_batch_get_timing": ([vp, C.c_int, C.POINTER(f64), C.POINTER(i64)], C.c_int),
}

SIGNATURES = {
    "pm_abi_version": ([], C.c_int),
    "pm_create": ([C.c_int, vp, C.POINTER(vp)], C.c_int),
    "pm_destroy": ([vp], C.c_int),
    "pm_last_error": ([vp], C.c_char_p),
    "pm_configure": ([vp, C.POINTER(pm_plan)], C.c_int),
    "pm_configure_static": ([vp, C.POINTER(pm_plan)], C.c_int),
    "pm_update_h": ([vp], C.c_int),
    "pm_update_e": ([vp], C.c_int),
    "pm_run": ([vp, i64, i32, C.c_uint32], C.c_int),
    "pm_dft_flush": ([vp, i64, i32], C.c_int),
    "pm_set_tuning": ([vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
    "pm_fused_active": ([vp], C.c_int),
    "pm_sync": ([vp], C.c_int),
    "pm_set_row_split": ([vp, C.c_int], C.c_int),
    "pm_row_split_active": ([vp], C.c_int),
    "pm_fused_shape": ([vp, C.POINTER(i32)], C.c_int),
    "pm_set_timing": ([vp, C.c_int], C.c_int),
    "pm_get_timing": ([vp, C.c_int, C.POINTER(f64), C.POINTER(i64)], C.c_int),
}

_LIB = None


class NativeError(RuntimeError):
    pass


def load(path: os.PathLike | None = None):
    """Load libpmfdtd once; raise if it is missing or of another ABI."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else Path(os.environ.get("PRISMO_AMD_LIB", LIB_PATH))
    if not p.exists():
        raise NativeError(
            f"{p} is missing: build it with `python -m prismo_amd.build_native` "
            "(the mi355x time step has no CPU fallback)"
        )
    lib = C.CDLL(str(p))
    for name, (args, res) in {**SIGNATURES, **BATCH_SIGNATURES, **MLP_SIGNATURES,
                              **GEOM_SIGNATURES}.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if lib.pm_abi_version() != ABI_VERSION:
        raise NativeError(
            f"libpmfdtd ABI {lib.pm_abi_version()} != binding ABI {ABI_VERSION}"
        )
    if path is None:
        _LIB = lib
    return lib


class Context:
    """One pm_ctx on a device/stream, with status codes mapped to exceptions."""

    def __init__(self, device: int = 0, stream: int | None = None):
        self.lib = load()
        h = vp()
        rc = self.lib.pm_create(device, vp(stream or 0), C.byref(h))
        if rc != 0:
            raise NativeError(f"pm_create failed ({rc})")
        self.h = h
        self._keep = []

    def _check(self, rc: int, what: str):
        if rc == 0:
            return
        msg = self.lib.pm_last_error(self.h).decode(errors="replace")
        kind = ERRORS.get(rc, f"HIP error {rc}")
        if rc in (-1, -4):
            raise ValueError(f"{what}: {kind}: {msg}")
        if rc == -2:
            raise NotImplementedError(f"{what}: {kind}: {msg}")
        raise NativeError(f"{what}: {kind}: {msg}")

    def configure(self, plan: pm_plan, keep=()):
        self._keep = list(keep)  # host arrays referenced by the plan
        self._check(self.lib.pm_configure(self.h, C.byref(plan)), "pm_configure")

    def configure_static(self, plan: pm_plan, keep=()):
        """pm_configure_static: the library's own per-device scratch instead
        of plan.scratch (which must be NULL); allocates no device memory."""
        self._keep = list(keep)
        self._check(self.lib.pm_configure_static(self.h, C.byref(plan)), "pm_configure_static")

    def update_h(self):
        self._check(self.lib.pm_update_h(self.h), "pm_update_h")

    def update_e(self):
        self._check(self.lib.pm_update_e(self.h), "pm_update_e")

    def run(self, step0: int, nsteps: int, flags: int):
        self._check(self.lib.pm_run(self.h, step0, nsteps, flags), "pm_run")

    def dft_flush(self, step_first: int = 0, rows: int = 0):
        """rows == 0: sum the samples pm_run has not summed yet; rows > 0:
        sum rows of X as the samples of steps [step_first, +rows)."""
        self._check(self.lib.pm_dft_flush(self.h, step_first, rows), "pm_dft_flush")

    def set_tuning(self, fuse=True, ty: int = 4, kchunk: int = 16, vec: int = 0,
                   order: int = 0):
        """fuse: False/0 unfused, True/1 one step per launch, 2 two steps
        per launch (ty 4, vec 0)."""
        f = int(fuse) if not isinstance(fuse, bool) else (1 if fuse else 0)
        self._check(self.lib.pm_set_tuning(self.h, f, ty, kchunk, vec, order), "pm_set_tuning")

    def fused_active(self) -> bool:
        return bool(self.lib.pm_fused_active(self.h))

    def set_row_split(self, segments: int = 0):
        """Fused-kernel rows in 1 or 2 workgroups (0: automatic)."""
        self._check(self.lib.pm_set_row_split(self.h, int(segments)), "pm_set_row_split")

    def fused_shape(self):
        """(vec, w, ty, xs) of the next one-step fused launch, or None."""
        out = (i32 * 4)()
        if not self.lib.pm_fused_shape(self.h, out):
            return None
        return tuple(int(v) for v in out)

    def row_split(self) -> int:
        """Row segments the next fused step takes (0 when not fused)."""
        return int(self.lib.pm_row_split_active(self.h))

    def fused_steps(self) -> int:
        """Steps per fused launch of a run without monitor sampling (0, 1, 2)."""
        return int(self.lib.pm_fused_active(self.h))

    def sync(self):
        self._check(self.lib.pm_sync(self.h), "pm_sync")

    def set_timing(self, on, stride: int = 1):
        """Time the kernels of every ``stride``-th step (off when ``on`` is false)."""
        self._check(self.lib.pm_set_timing(self.h, max(1, int(stride)) if on else 0),
                    "pm_set_timing")

    def timing(self, which: int):
        ms = f64()
        n = i64()
        self._check(self.lib.pm_get_timing(self.h, which, C.byref(ms), C.byref(n)), "pm_get_timing")
        return ms.value, n.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.pm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchContext:
    """One pm_batch on a device/stream (include/pmbatch.h)."""

    def __init__(self, device: int = 0, stream: int | None = None):
        self.lib = load()
        h = vp()
        rc = self.lib.pm_batch_create(device, vp(stream or 0), C.byref(h))
        if rc != 0:
            raise NativeError(f"pm_batch_create failed ({rc})")
        self.h = h
        self._keep = []

    def _check(self, rc: int, what: str):
        if rc == 0:
            return
        msg = self.lib.pm_batch_last_error(self.h).decode(errors="replace")
        kind = ERRORS.get(rc, f"HIP error {rc}")
        if rc in (-1, -4):
            raise ValueError(f"{what}: {kind}: {msg}")
        if rc == -2:
            raise NotImplementedError(f"{what}: {kind}: {msg}")
        raise NativeError(f"{what}: {kind}: {msg}")

    def workspace_bytes(self, plan: pm_batch_plan) -> int:
        n = i64()
        rc = self.lib.pm_batch_workspace_bytes(C.byref(plan), C.byref(n))
        if rc != 0:
            raise ValueError("pm_batch_workspace_bytes: invalid plan")
        return n.value

    def configure(self, plan: pm_batch_plan, keep=()):
        self._keep = list(keep)
        self._check(self.lib.pm_batch_configure(self.h, C.byref(plan)), "pm_batch_configure")

    def run(self, nsteps: int, graph: bool = False):
        self._check(self.lib.pm_batch_run(self.h, nsteps, RUN_GRAPH if graph else 0),
                    "pm_batch_run")

    def flush(self):
        self._check(self.lib.pm_batch_flush(self.h), "pm_batch_flush")

    def parity(self) -> int:
        return int(self.lib.pm_batch_parity(self.h))

    def steps(self) -> int:
        return int(self.lib.pm_batch_steps(self.h))

    def sync(self):
        self._check(self.lib.pm_batch_sync(self.h), "pm_batch_sync")

    def set_tuning(self, waves: int = 4, xcd: int = 1):
        """Waves per workgroup (1, 2, 4, 8) and XCD-grouped order (0/1)."""
        self._check(self.lib.pm_batch_set_tuning(self.h, int(waves), int(xcd)),
                    "pm_batch_set_tuning")

    def design_status(self, n: int):
        """Per design, bit 0: an Ez value was Inf or NaN (waits for the stream)."""
        import numpy as np

        out = (i32 * n)()
        self._check(self.lib.pm_batch_design_status(self.h, out, n), "pm_batch_design_status")
        return np.array(out[:], dtype=np.int32)

    def set_timing(self, on: bool):
        self._check(self.lib.pm_batch_set_timing(self.h, 1 if on else 0), "pm_batch_set_timing")

    def timing(self, which: int):
        ms = f64()
        n = i64()
        self._check(self.lib.pm_batch_get_timing(self.h, which, C.byref(ms), C.byref(n)),
                    "pm_batch_get_timing")
        return ms.value, n.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.pm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
