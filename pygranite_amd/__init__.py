"""pygranite_amd: pygranite-compatible trajectory engine for AMD Instinct MI355X (gfx950).

``from pygranite_amd import pygranite`` (or ``import pygranite_amd as pygranite``) gives the
reference's Python API: IntegratorSettings, DataLoader, TrajectorySet, TrajectoryIntegrator and
the enums. The compute path is the HIP library ``libgranite_hip.so`` behind the C ABI declared
in ``include/granite_hip.h``; there is no CPU fallback -- importing fails loudly when the
native module has not been built (``python -m pygranite_amd.build``).
"""
from __future__ import annotations

import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "libgranite_hip.so")

try:
    from . import pygranite  # noqa: F401  (the pybind11 extension)
except ImportError as e:  # pragma: no cover - exercised only when unbuilt
    raise ImportError(
        "pygranite_amd native module is not built: run `python -m pygranite_amd.build` "
        f"(hipcc, gfx950). Original error: {e}") from e

from .pygranite import *  # noqa: F401,F403,E402

__all__ = [n for n in dir(pygranite) if not n.startswith("_")]
