"""wslda_amd -- MI355X-native time-dependent W-SLDA step (HIP, gfx950).

The compute path is libwslda_td.so (hand-written HIP kernels behind a C ABI,
include/wslda_td.h); wslda_amd.td binds it for Python callers.
"""
from . import td  # noqa: F401

__all__ = ["td"]
