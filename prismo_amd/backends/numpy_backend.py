"""Host NumPy arrays behind the ``Backend`` interface.

Used for host-side set-up arrays (material maps, index tables).  The time
step itself never runs on this backend in this package: ``MaxwellUpdater``
refuses it and names the 'mi355x' backend instead.  Method set mirrors
src/prismo/backends/numpy_backend.py:15-159.
"""

import numpy as np

from .base import Backend


class NumPyBackend(Backend):
    name = property(lambda self: "numpy")
    is_gpu = property(lambda self: False)

    def zeros(self, shape, dtype=None):
        return np.zeros(shape, dtype=dtype or np.float64)

    def ones(self, shape, dtype=None):
        return np.ones(shape, dtype=dtype or np.float64)

    def empty(self, shape, dtype=None):
        return np.empty(shape, dtype=dtype or np.float64)

    def array(self, data, dtype=None):
        return np.array(data, dtype=dtype)

    def asarray(self, data, dtype=None):
        return np.asarray(data, dtype=dtype)

    def to_numpy(self, array):
        return np.asarray(array)

    def copy(self, array):
        return np.copy(array)

    sqrt = staticmethod(np.sqrt)
    exp = staticmethod(np.exp)
    sin = staticmethod(np.sin)
    cos = staticmethod(np.cos)
    abs = staticmethod(np.abs)

    def sum(self, array, axis=None):
        return np.sum(array, axis=axis)

    def max(self, array, axis=None):
        return np.max(array, axis=axis)

    def min(self, array, axis=None):
        return np.min(array, axis=axis)

    def mean(self, array, axis=None):
        return np.mean(array, axis=axis)

    def fft(self, array, axis=-1):
        return np.fft.fft(array, axis=axis)

    def ifft(self, array, axis=-1):
        return np.fft.ifft(array, axis=axis)

    def fft2(self, array):
        return np.fft.fft2(array)

    def ifft2(self, array):
        return np.fft.ifft2(array)

    def dot(self, a, b):
        return np.dot(a, b)

    def matmul(self, a, b):
        return np.matmul(a, b)

    def where(self, condition, x, y):
        return np.where(condition, x, y)

    def synchronize(self):
        return None

    def get_memory_info(self):
        import psutil

        mi = psutil.Process().memory_info()
        return {
            "backend": "numpy",
            "device": "CPU",
            "rss_bytes": mi.rss,
            "rss_mb": mi.rss / 2**20,
            "vms_bytes": mi.vms,
            "vms_mb": mi.vms / 2**20,
        }

    float32 = property(lambda self: np.float32)
    float64 = property(lambda self: np.float64)
    complex64 = property(lambda self: np.complex64)
    complex128 = property(lambda self: np.complex128)
    int32 = property(lambda self: np.int32)
    int64 = property(lambda self: np.int64)
    pi = property(lambda self: np.pi)
