"""MI355X-native two-loop SPH hot path (density + Newton h, force/second
derivative loop, neighbour search) behind the reference engine's interface.

The compute path is libsph_hip.so (hand-written HIP for gfx950, C-ABI in
include/sph_hip.h); this package holds the Python host mirror of the
reference interface (``Engine``/``iterate``), the reference enumerators
(``consts``) and seeded workload generators (``problems``).
"""
from . import consts, problems  # noqa: F401

__all__ = ["consts", "problems", "Engine", "iterate", "sts_coeffs"]


def __getattr__(name):
    # the HIP-backed symbols load libsph_hip.so on first use
    if name in ("Engine", "iterate", "sts_coeffs"):
        from . import engine
        return getattr(engine, name)
    raise AttributeError(name)
