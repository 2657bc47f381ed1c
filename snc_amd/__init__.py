"""MI355X (gfx950) backend for SNC's event-driven spiking path.

Hand-written HIP kernels behind a C ABI (include/snc_hip.h); this package is
the host-side mirror of the reference's runtime and trainer interfaces.
"""
from . import _lib  # noqa: F401
from . import grow, structure, train  # noqa: F401  (bind trainer, MLP and structure entry points)
from .graph import SNNGraph  # noqa: F401
from .runtime import Backend, ForwardPlan, ForwardResult, HipGraph, LIFParams, available, forward  # noqa: F401
