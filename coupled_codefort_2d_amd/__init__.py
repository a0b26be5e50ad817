"""MI355X-native (gfx950) FUGU-2D per-step hot path.

The product is the C-ABI library libfugu2d_hip.so (include/fugu2d.h) built from csrc/,
plus the Fortran drop-in interface modules under fortran/.  This Python package only
builds the library (build.py) and binds it for Python callers (abi.py).
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
