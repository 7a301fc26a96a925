"""MI355X (gfx950) implementation of PedalSky's sky + cloud transport path.

The compute lives in ``lib/libpedalsky_hip.so`` (hand-written HIP kernels
behind the C ABI of ``include/pedalsky_hip.h``). This package holds the
Python side of that boundary: the ctypes mirror (``abi``), a context handle
(``renderer``), scene/workload construction (``scene``) and the host mirror
of the reference's RenderSubsystem interface (``subsystem``).
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
