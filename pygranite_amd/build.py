"""Builds the native parts of pygranite_amd in-tree.

* ``libgranite_hip.so`` -- gfx950 HIP kernels (csrc/gr_k_*.hip) + the C ABI of include/granite_hip.h
  (hipcc, --offload-arch=gfx950).
* ``pygranite.<ext>`` -- the pybind11 ``pygranite`` module (g++), linked against
  libgranite_hip.so with an $ORIGIN rpath.

Run ``python -m pygranite_amd.build`` (or ``__graft_entry__.build()``). hipcc cross-compiles
without a GPU. Every call rebuilds.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("GRANITE_OFFLOAD_ARCH", "gfx950")

LIB = os.path.join(PKG, "libgranite_hip.so")
EXT = os.path.join(PKG, "pygranite" + sysconfig.get_config_var("EXT_SUFFIX"))

# Device code: no FP contraction (the FP32 form performs exactly the IEEE operations of the
# step), correctly rounded FP32 divide/sqrt, and no packed-FP32 instructions.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function",
]
# kernel translation units (compiled in parallel): the single-job step kernels per storage type
# (the FP32 one also holds their launcher), the batched kernel's three builds and its
# per-variant kernels (2D, 3D), the support kernels, the allocation-free entry point's kernels and
# the field-tile kernels; all share gr_step.h
HIP_SOURCES = ("gr_k_adv_f32.hip", "gr_k_adv_f16.hip", "gr_k_adv_bf16.hip", "gr_k_batch.hip",
               "gr_k_batch_w4.hip", "gr_k_batch_narrow.hip", "gr_k_batch_uni2.hip", "gr_k_batch_uni3.hip",
               "gr_k_support.hip", "gr_k_noalloc.hip", "gr_k_adv_tile.hip")
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
              "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r


def build(verbose: bool = False) -> None:
    os.makedirs(BUILD, exist_ok=True)
    inc = [f"-I{INCLUDE}", f"-I{CSRC}"]
    objs = []
    jobs = []
    for src in HIP_SOURCES:
        o = os.path.join(BUILD, src.replace(".hip", ".o"))
        jobs.append([HIPCC, *HIP_FLAGS, *inc, "-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    for src in ("gr_capi.cpp", "gr_expr.cpp"):
        o = os.path.join(BUILD, src.replace(".cpp", ".o"))
        jobs.append(["g++", *HOST_FLAGS, *inc, "-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=min(10, (os.cpu_count() or 4) + 2)) as ex:
        list(ex.map(_run, jobs))
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", LIB + ".tmp"])
    os.replace(LIB + ".tmp", LIB)

    import pybind11
    py_inc = sysconfig.get_paths()["include"]
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
          f"-I{pybind11.get_include()}", f"-I{py_inc}", *inc,
          os.path.join(CSRC, "pygranite_module.cpp"),
          f"-L{PKG}", "-lgranite_hip", "-Wl,-rpath,$ORIGIN", "-o", EXT + ".tmp"])
    os.replace(EXT + ".tmp", EXT)
    if verbose:
        print("built", LIB)
        print("built", EXT)


if __name__ == "__main__":
    build(verbose=True)
