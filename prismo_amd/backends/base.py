"""Array-backend interface (the reference's plugin contract).

Same abstract surface as src/prismo/backends/base.py:14-219: properties
``name``/``is_gpu``, allocation (``zeros``/``ones``/``empty``/``array``/
``asarray``/``copy``/``to_numpy``), elementwise math, reductions, FFTs,
``dot``/``matmul``/``where``, ``synchronize``, ``get_memory_info``, dtype
properties and ``pi``.
"""

from abc import ABC, abstractmethod
from typing import Any


class Backend(ABC):
    @property
    @abstractmethod
    def name(self) -> str: ...

    @property
    @abstractmethod
    def is_gpu(self) -> bool: ...

    @abstractmethod
    def zeros(self, shape, dtype: Any = None): ...

    @abstractmethod
    def ones(self, shape, dtype: Any = None): ...

    @abstractmethod
    def empty(self, shape, dtype: Any = None): ...

    @abstractmethod
    def array(self, data, dtype: Any = None): ...

    @abstractmethod
    def asarray(self, data, dtype: Any = None): ...

    @abstractmethod
    def to_numpy(self, array): ...

    @abstractmethod
    def copy(self, array): ...

    @abstractmethod
    def sqrt(self, array): ...

    @abstractmethod
    def exp(self, array): ...

    @abstractmethod
    def sin(self, array): ...

    @abstractmethod
    def cos(self, array): ...

    @abstractmethod
    def abs(self, array): ...

    @abstractmethod
    def sum(self, array, axis=None): ...

    @abstractmethod
    def max(self, array, axis=None): ...

    @abstractmethod
    def min(self, array, axis=None): ...

    @abstractmethod
    def mean(self, array, axis=None): ...

    @abstractmethod
    def fft(self, array, axis: int = -1): ...

    @abstractmethod
    def ifft(self, array, axis: int = -1): ...

    @abstractmethod
    def fft2(self, array): ...

    @abstractmethod
    def ifft2(self, array): ...

    @abstractmethod
    def dot(self, a, b): ...

    @abstractmethod
    def matmul(self, a, b): ...

    @abstractmethod
    def where(self, condition, x, y): ...

    @abstractmethod
    def synchronize(self) -> None: ...

    @abstractmethod
    def get_memory_info(self) -> dict: ...

    @property
    @abstractmethod
    def float32(self): ...

    @property
    @abstractmethod
    def float64(self): ...

    @property
    @abstractmethod
    def complex64(self): ...

    @property
    @abstractmethod
    def complex128(self): ...

    @property
    @abstractmethod
    def int32(self): ...

    @property
    @abstractmethod
    def int64(self): ...

    @property
    @abstractmethod
    def pi(self) -> float: ...

    def __repr__(self) -> str:
        dev = "GPU" if self.is_gpu else "CPU"
        return f"{self.__class__.__name__}(name='{self.name}', device={dev})"
