"""hann_amd: MI355X (gfx950) implementation of hann's batch-distance and
PQIVF hot path behind a C-ABI (include/hann_amd.h).

`core` mirrors the reference's core package (metrics, batch distances),
`pqivf` mirrors pqivf.Index. Both call libhann_amd.so through ctypes; there
is no CPU fallback.
"""
from . import _lib  # noqa: F401
from ._lib import HannError  # noqa: F401

__all__ = ["core", "pqivf", "HannError"]


def __getattr__(name):
    if name in ("core", "pqivf", "serving"):
        import importlib
        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)
