import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libpmfdtd.so")
    config.addinivalue_line("markers", "slow: longer CPU/GPU cases")


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def mi355x():
    """The mi355x backend; a GPU test fails (not skips) when the HIP library
    cannot be loaded, so a missing extension is never silently green."""
    if not gpu_available():
        pytest.skip("no HIP device visible")
    from prismo_amd import _native, set_backend

    _native.load()
    return set_backend("mi355x")


def rel_l2(a, b):
    import numpy as np

    a = np.asarray(a, dtype=np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64)
    b = np.asarray(b, dtype=a.dtype)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return num / den if den > 0 else num
