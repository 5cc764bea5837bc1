"""Waveguide eigenmodes (src/prismo/modes/__init__.py)."""

from .solver import ModeSolver, WaveguideMode

__all__ = ["ModeSolver", "WaveguideMode"]
