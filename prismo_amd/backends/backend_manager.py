"""Backend registry: ``set_backend`` / ``get_backend`` / ``list_available_backends``.

Same contract as src/prismo/backends/backend_manager.py:33-241 (a module-level
current backend, ``ValueError`` naming the available backends for an unknown
or unavailable name), with 'mi355x' registered.  'cupy' and 'metal' are not
part of this package and report as unavailable.
"""

from __future__ import annotations

import warnings
from typing import Optional

from .base import Backend
from .numpy_backend import NumPyBackend

_CURRENT_BACKEND: Optional[Backend] = None


def _mi355x_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


def list_available_backends() -> list:
    out = ["numpy"]
    if _mi355x_available():
        out.append("mi355x")
    return out


def get_backend_info() -> dict:
    info = {
        "available_backends": list_available_backends(),
        "current_backend": _CURRENT_BACKEND.name if _CURRENT_BACKEND else None,
        "numpy_available": True,
        "mi355x_available": _mi355x_available(),
    }
    if info["mi355x_available"]:
        import torch

        info["num_devices"] = torch.cuda.device_count()
        info["hip_version"] = torch.version.hip
        info["devices"] = [
            {"id": i, "name": torch.cuda.get_device_name(i)}
            for i in range(info["num_devices"])
        ]
    return info


def set_backend(backend: str, device_id: int = 0) -> Backend:
    """Select the global backend ('numpy' or 'mi355x')."""
    global _CURRENT_BACKEND
    name = backend.lower()
    if name == "numpy":
        _CURRENT_BACKEND = NumPyBackend()
    elif name == "mi355x":
        if not _mi355x_available():
            raise ValueError(
                "mi355x backend requested but no ROCm GPU is visible to torch"
            )
        from .mi355x_backend import MI355XBackend

        _CURRENT_BACKEND = MI355XBackend(device_id=device_id)
    else:
        raise ValueError(
            f"Unknown backend '{backend}'. Available backends: "
            f"{list_available_backends()}"
        )
    return _CURRENT_BACKEND


def get_backend(backend: Optional[str] = None, device_id: int = 0) -> Backend:
    """Named backend, else the current one, else mi355x when a GPU is visible."""
    if backend is not None:
        return set_backend(backend, device_id)
    if _CURRENT_BACKEND is not None:
        return _CURRENT_BACKEND
    if _mi355x_available():
        warnings.warn(
            "No backend specified. Auto-selecting the mi355x (GPU) backend.",
            UserWarning,
            stacklevel=2,
        )
        return set_backend("mi355x", device_id)
    return set_backend("numpy")


def auto_select_backend() -> Backend:
    return get_backend()


def resolve_backend(backend) -> Backend:
    """Constructor argument handling shared by every class that allocates
    (str name, Backend instance, or None; TypeError otherwise — the rule of
    src/prismo/core/solver.py:51-59)."""
    if isinstance(backend, str):
        return get_backend(backend)
    if isinstance(backend, Backend):
        return backend
    if backend is None:
        return get_backend()
    raise TypeError("backend must be a Backend instance or string name")
