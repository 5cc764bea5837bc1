from .base import Monitor
from .dft import DFTMonitor
from .field import FieldMonitor

__all__ = ["Monitor", "FieldMonitor", "DFTMonitor"]
