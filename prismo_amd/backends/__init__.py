"""Backend plugin layer (src/prismo/backends/__init__.py) with 'mi355x'."""

from .backend_manager import (
    get_backend,
    get_backend_info,
    list_available_backends,
    resolve_backend,
    set_backend,
)
from .base import Backend
from .numpy_backend import NumPyBackend

__all__ = ["Backend", "NumPyBackend", "get_backend", "set_backend",
           "list_available_backends", "get_backend_info", "resolve_backend"]
