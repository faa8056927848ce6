"""saruman_amd -- MI355X-native semiglobal Needleman-Wunsch alignment with traceback,
a drop-in for SARUMAN's GPU module (src/cuda-functions.h).

The compute path is libsaruman_hip.so (hand-written HIP for gfx950, C ABI in
include/saruman_hip.h).  This package is the Python host side: ctypes bindings,
2-bit packing, batch helpers.  It has no CPU fallback: without the built library
`load()` raises.
"""
from __future__ import annotations

from . import _lib, testing  # noqa: F401
from ._lib import load  # noqa: F401
from .batch import AlignResult, Batch, Context, Job, align_batch, pack_codes  # noqa: F401

__all__ = ["load", "Batch", "AlignResult", "Context", "Job", "align_batch", "pack_codes"]
