"""pygranite -- the reference's import name for the MI355X trajectory engine.

``import pygranite`` gives the pybind11 extension module itself (built from
pygranite_amd/csrc/pygranite_module.cpp, module name ``pygranite``), with the same enums,
IntegratorSettings, DataLoader, TrajectorySet and TrajectoryIntegrator as the reference
(source/pygranite.cpp:3-152). The extension lives in the pygranite_amd package beside
libgranite_hip.so; this package only puts it under the reference's name.
"""
import sys as _sys

from pygranite_amd import native as _native

_sys.modules[__name__] = _native()
