"""MI355X (gfx950) implementation of PedalSky's sky + cloud transport path.

The compute lives in ``lib/libpedalsky_hip.so`` (hand-written HIP kernels
behind the C ABI of ``include/pedalsky_hip.h``). This package holds the
Python side of that boundary: the ctypes mirror (``abi``), the context
handle that mirrors the reference's RenderSubsystem prepare / dispatch calls
(``renderer``), scene and workload construction (``scene``), many scenes per
launch (``batch``), the autograd binding of the cloud adjoint (``autograd``),
the atmosphere calibration (``calibration``), the roofline cost table
(``cost_model``) and the in-tree build (``build``).
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
