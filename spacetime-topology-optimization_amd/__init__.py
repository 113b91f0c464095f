"""MI355X-native solve-and-sensitivity path of sttopt (space-time topology optimization).

Modules mirror the reference's hot-path modules and call signatures:
`torch_fem`, `torch_mg`, `torch_solve`, `compliance`, `sensitivity`; `fem` holds the
mesh setup, `virtual_heat` the scalar (heat) solve, `workloads` the synthetic problems.
Every solve runs in the gfx950 HIP library `libsttopt_hip.so` (C ABI:
include/sttopt_hip.h), bound with ctypes in `_lib`.
"""

from . import _lib, fem  # noqa: F401

__all__ = [
    "compliance",
    "fem",
    "sensitivity",
    "torch_fem",
    "torch_mg",
    "torch_solve",
    "virtual_heat",
    "workloads",
]


def __getattr__(name):
    if name in __all__:
        import importlib

        return importlib.import_module(f"{__name__}.{name}")
    raise AttributeError(name)
