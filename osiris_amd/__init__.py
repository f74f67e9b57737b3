"""osiris_amd -- MI355X (gfx950) per-cell Newton hot path of the Osiris
compositional dual-porosity reservoir simulator.

The compute lives in ``libosiris_amd.so`` (hand-written HIP kernels behind a
C ABI, declared in ``include/osiris_amd.h``).  This module is the thin host
side: it loads that library with ctypes and passes it device pointers of
torch tensors (torch is used for device memory and streams only).  There is
no CPU fallback: if the library is missing, importing the runtime raises.
"""
from __future__ import annotations

from . import synth  # noqa: F401  (pure numpy, no GPU)

__all__ = ["synth", "runtime"]
