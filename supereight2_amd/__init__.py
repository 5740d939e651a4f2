"""MI355X-native single-resolution TSDF pipeline (allocation, update, raycast, ICP).

The product is libse8.so (HIP kernels for gfx950 behind the C-ABI in include/se8.h) and the
header-only C++ facade include/se8/supereight.hpp.  This package holds the HIP sources
(csrc/), the in-tree build (build.py), the ctypes binding used by the Python check harness
(capi.py, se.py) and the seeded synthetic workloads (synth.py).
"""
from . import capi, synth  # noqa: F401

__all__ = ["capi", "synth"]
