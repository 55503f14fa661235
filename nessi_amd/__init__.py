"""nessi_amd: MI355X-native Kadanoff-Baym time stepping behind the cntr:: interface.

The compute path is libcntrx.so (HIP, gfx950, complex FP64) reached through its C ABI
(include/cntrx.h); `nessi_amd.cntr` mirrors the reference's cntr:: functions.
"""
from . import _lib  # noqa: F401

__all__ = ["cntr", "lattice", "batch_server", "multiorbital", "workloads"]


def __getattr__(name):
    if name in __all__:
        import importlib
        return importlib.import_module("." + name, __name__)
    raise AttributeError(name)
