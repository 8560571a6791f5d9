"""untangling26_amd: gfx950 (MI355X) implementation of the Untangling26 per-step hot path.

The compute path is libuntangle_hip.so (hand-written HIP kernels behind the C
ABI of include/untangle.h).  Python modules here mirror the reference's
Solver/ interfaces on top of that ABI; they hold no compute fallback.
"""
from ._lib import LIB_PATH, lib  # noqa: F401

__all__ = ["LIB_PATH", "lib"]
