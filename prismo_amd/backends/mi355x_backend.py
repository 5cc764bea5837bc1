"""The 'mi355x' backend: torch-ROCm device arrays plus the field layout the
HIP time-step library expects.

Generic arrays (``zeros``, ``asarray``, ...) are plain torch tensors on the
selected device, default float64 like the reference's CuPy backend
(src/prismo/backends/cupy_backend.py:54-72).  Field components are different:
``field_zeros`` allocates one zero-filled x-fastest buffer of shape
``(NZ, NY, PX)`` per component, shared pitch ``PX`` for all six, and returns a
permuted view whose shape and ``[i, j, k]`` indexing are the reference's
(src/prismo/core/grid.py:137-176).  So ``fields['Ez'][2, 2] = 1.0`` and
``field[np.ix_(...)] += a`` keep working, while the kernels see coalesced rows
along x.  Cells of the padded box outside a component's extent stay zero;
kernels read them as the zero field beyond the domain edge and never write
them.
"""

from __future__ import annotations

import numpy as np

from .base import Backend

_PITCH_ALIGN = 64  # elements; keeps each x-row start 256/512-byte aligned


def pitch_for(nx: int) -> int:
    return ((nx + _PITCH_ALIGN - 1) // _PITCH_ALIGN) * _PITCH_ALIGN


def _torch():
    import torch

    return torch


_NP2TORCH = None


def torch_dtype(dtype):
    """Map numpy / python / torch dtypes to a torch dtype (default float64)."""
    torch = _torch()
    global _NP2TORCH
    if _NP2TORCH is None:
        _NP2TORCH = {
            np.dtype(np.float64): torch.float64,
            np.dtype(np.float32): torch.float32,
            np.dtype(np.float16): torch.float16,
            np.dtype(np.complex128): torch.complex128,
            np.dtype(np.complex64): torch.complex64,
            np.dtype(np.int32): torch.int32,
            np.dtype(np.int64): torch.int64,
            np.dtype(np.bool_): torch.bool,
            np.dtype(np.uint8): torch.uint8,
        }
    if dtype is None:
        return torch.float64
    if isinstance(dtype, torch.dtype):
        return dtype
    return _NP2TORCH[np.dtype(dtype)]


class MI355XBackend(Backend):
    """torch-ROCm arrays on one MI355X; the stencil runs in libpmfdtd."""

    def __init__(self, device_id: int = 0):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError(
                "mi355x backend requested but no ROCm GPU is visible to torch"
            )
        self.device_id = device_id
        self.device = torch.device("cuda", device_id)
        torch.cuda.set_device(self.device)

    name = property(lambda self: "mi355x")
    is_gpu = property(lambda self: True)

    # -- allocation --------------------------------------------------------
    def zeros(self, shape, dtype=None):
        return _torch().zeros(tuple(shape), dtype=torch_dtype(dtype), device=self.device)

    def ones(self, shape, dtype=None):
        return _torch().ones(tuple(shape), dtype=torch_dtype(dtype), device=self.device)

    def empty(self, shape, dtype=None):
        return _torch().empty(tuple(shape), dtype=torch_dtype(dtype), device=self.device)

    def array(self, data, dtype=None):
        torch = _torch()
        if isinstance(data, torch.Tensor):
            t = data.to(self.device).clone()
            return t if dtype is None else t.to(torch_dtype(dtype))
        a = np.array(data, dtype=dtype)
        return torch.from_numpy(a).to(self.device)

    def asarray(self, data, dtype=None):
        torch = _torch()
        if isinstance(data, torch.Tensor):
            t = data.to(self.device)
            return t if dtype is None else t.to(torch_dtype(dtype))
        a = np.ascontiguousarray(np.asarray(data, dtype=dtype))
        return torch.from_numpy(a).to(self.device)

    def to_numpy(self, array):
        torch = _torch()
        if isinstance(array, torch.Tensor):
            return np.ascontiguousarray(array.detach().cpu().numpy())
        return np.asarray(array)

    def copy(self, array):
        return array.clone()

    def field_zeros(self, grid, component: str, dtype=None):
        """Zeroed component in the padded x-fastest layout (reference-shaped view)."""
        torch = _torch()
        shape = grid.get_field_shape(component)
        nx, ny, nz = grid.dimensions
        buf = torch.zeros(
            (nz, ny, pitch_for(nx)), dtype=torch_dtype(dtype), device=self.device
        )
        if grid.is_2d:
            return buf[0, : shape[1], : shape[0]].permute(1, 0)
        return buf[: shape[2], : shape[1], : shape[0]].permute(2, 1, 0)

    def field_set(self, grid, dtype=None) -> dict:
        """All six components in one zeroed allocation, Ex Ey Ez Hx Hy Hz at a
        constant stride (the layout the fused step kernel addresses), each
        returned as its reference-shaped view."""
        torch = _torch()
        nx, ny, nz = grid.dimensions
        arena = torch.zeros((6, nz, ny, pitch_for(nx)), dtype=torch_dtype(dtype),
                            device=self.device)
        out = {}
        for n, c in enumerate(("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")):
            shape = grid.get_field_shape(c)
            buf = arena[n]
            if grid.is_2d:
                out[c] = buf[0, : shape[1], : shape[0]].permute(1, 0)
            else:
                out[c] = buf[: shape[2], : shape[1], : shape[0]].permute(2, 1, 0)
        return out

    # -- math ------------------------------------------------------------------
    def sqrt(self, a):
        return _torch().sqrt(a)

    def exp(self, a):
        return _torch().exp(a)

    def sin(self, a):
        return _torch().sin(a)

    def cos(self, a):
        return _torch().cos(a)

    def abs(self, a):
        return _torch().abs(a)

    def sum(self, a, axis=None):
        return a.sum() if axis is None else a.sum(dim=axis)

    def max(self, a, axis=None):
        return a.max() if axis is None else a.amax(dim=axis)

    def min(self, a, axis=None):
        return a.min() if axis is None else a.amin(dim=axis)

    def mean(self, a, axis=None):
        return a.mean() if axis is None else a.mean(dim=axis)

    def fft(self, a, axis=-1):
        return _torch().fft.fft(a, dim=axis)

    def ifft(self, a, axis=-1):
        return _torch().fft.ifft(a, dim=axis)

    def fft2(self, a):
        return _torch().fft.fft2(a)

    def ifft2(self, a):
        return _torch().fft.ifft2(a)

    def dot(self, a, b):
        torch = _torch()
        if a.dim() == 1 and b.dim() == 1:
            return torch.dot(a, b)
        return torch.matmul(a, b)

    def matmul(self, a, b):
        return _torch().matmul(a, b)

    def where(self, condition, x, y):
        return _torch().where(condition, x, y)

    def synchronize(self):
        _torch().cuda.synchronize(self.device)

    def get_memory_info(self):
        torch = _torch()
        free, total = torch.cuda.mem_get_info(self.device)
        return {
            "backend": "mi355x",
            "device": torch.cuda.get_device_name(self.device),
            "device_id": self.device_id,
            "total_bytes": total,
            "free_bytes": free,
            "used_bytes": total - free,
            "allocated_bytes": torch.cuda.memory_allocated(self.device),
        }

    float32 = property(lambda self: _torch().float32)
    float64 = property(lambda self: _torch().float64)
    complex64 = property(lambda self: _torch().complex64)
    complex128 = property(lambda self: _torch().complex128)
    int32 = property(lambda self: _torch().int32)
    int64 = property(lambda self: _torch().int64)
    pi = property(lambda self: float(np.pi))
