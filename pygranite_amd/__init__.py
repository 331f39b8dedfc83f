"""pygranite_amd: pygranite-compatible trajectory engine for AMD Instinct MI355X (gfx950).

``import pygranite`` (the top-level package installed beside this one), ``from pygranite_amd
import pygranite`` or ``import pygranite_amd as pygranite`` give the reference's Python API:
IntegratorSettings, DataLoader, TrajectorySet, TrajectoryIntegrator and the enums. The compute
path is the HIP library ``libgranite_hip.so`` behind the C ABI declared in
``include/granite_hip.h``; there is no CPU fallback.

The native module is loaded on first use of one of its names, so that the pure-Python parts
(``build``, ``synthetic``, ``capi``'s declarations) import on a tree that has not been built
yet. Using an API name there fails loudly with the build command.
"""
from __future__ import annotations

import importlib
import importlib.util
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "libgranite_hip.so")
BUILD_HINT = ("pygranite_amd native module is not built: run `python -m pygranite_amd.build` "
              "(hipcc, gfx950) or `pip install .`")

_native = None


def native():
    """The pybind11 extension ``pygranite`` (imported once; raises ImportError if unbuilt)."""
    global _native
    if _native is None:
        try:
            _native = importlib.import_module(".pygranite", __name__)
        except ImportError as e:
            raise ImportError(f"{BUILD_HINT}. Original error: {e}") from e
    return _native


def __getattr__(name: str):
    # PEP 562: API names (and the extension itself) resolve through the native module
    if name.startswith("__"):
        raise AttributeError(name)
    if name == "pygranite":
        return native()
    # `from pygranite_amd import build` asks for the attribute before importing the submodule:
    # pure-Python submodules must resolve without the native module
    if importlib.util.find_spec(f"{__name__}.{name}") is not None:
        return importlib.import_module(f".{name}", __name__)
    mod = native()
    try:
        return getattr(mod, name)
    except AttributeError:
        raise AttributeError(f"module 'pygranite_amd' has no attribute {name!r}") from None


def __dir__():
    names = set(globals())
    try:
        names.update(n for n in dir(native()) if not n.startswith("_"))
    except ImportError:
        pass
    return sorted(names)
