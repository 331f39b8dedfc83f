"""ORACLE -- test infrastructure only.

CPU restatement of pygranite's trajectory integrator (FP32 and FP64), loaded through ctypes
from ``oracle/liboracle.so`` (built by ``oracle/Makefile``). Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import this package,
and only as the checker or the timed CPU baseline. The product never routes through it.

``run()`` takes the reference's ``IntegratorSettings`` fields as a dict (same names), the full
stack of wind slices a ``DataLoader`` would hand out, and the start positions, and returns the
outputs the reference's ``TrajectorySet`` would hold, plus the step at which each particle
stopped (``abort_step``, -1 while alive).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

# enum values of source/granite/IntegratorSettings.hpp:16-76
SPACE = {"Space3D": 0, "Space2D": 1}
INTEGRATOR = {"ExplicitEuler": 0, "ClassicRungeKutta": 1}
BORDER = {"Block": 0, "LoopX": 1, "LoopY": 2, "LoopXY": 3}
CURVATURE = {"Off": 0, "TotalCurvature": 1, "FastTotalCurvature": 2,
             "IndividualAndTotalCurvature": 3}
ABORT_MODE = {"Time": 0, "Length": 1, "FitLength": 2}
MODE3 = {"Off": 0, "Constant": 1, "Dynamic": 2}
WINDFIELD = {"Constant": 0, "Dynamic": 1}
WEIGHTS = {"exact": 0, "q8": 1, "q8_round": 1, "q8_trunc": 2}

# ASTNodeType numbers of source/granite/Language.hpp:17-73 (+ 40 = abs extension)
NODE_F0, NODE_C0 = 5, 13


class _Args(ctypes.Structure):
    _fields_ = [
        ("real64", ctypes.c_int32),
        ("dim", ctypes.c_int32), ("integrator", ctypes.c_int32), ("border", ctypes.c_int32),
        ("curvature", ctypes.c_int32), ("abort_mode", ctypes.c_int32),
        ("uplift_mode", ctypes.c_int32), ("windfield_mode", ctypes.c_int32),
        ("constants_mode", ctypes.c_int32), ("volume_mode", ctypes.c_int32),
        ("reverse", ctypes.c_int32), ("weight_mode", ctypes.c_int32),
        ("time64", ctypes.c_int32), ("kahan", ctypes.c_int32),
        ("sphere_radius", ctypes.c_float), ("dt", ctypes.c_float), ("dtd", ctypes.c_float),
        ("max_length", ctypes.c_float), ("max_sim_time", ctypes.c_float),
        ("grid_scale", ctypes.c_float * 3), ("offset", ctypes.c_float * 3),
        ("save_interval", ctypes.c_uint64),
        ("min_alive", ctypes.c_int32), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32), ("nslices", ctypes.c_int32),
        ("nexpr", ctypes.c_int32), ("nconst", ctypes.c_int32), ("nvol", ctypes.c_int32),
        ("nthreads", ctypes.c_int32), ("allow_abs", ctypes.c_int32),
        ("exprs", ctypes.c_char_p * 4), ("const_names", ctypes.c_char_p * 4),
        ("vol_names", ctypes.c_char_p * 8),
        ("wind", ctypes.c_void_p), ("topo", ctypes.c_void_p), ("uplift", ctypes.c_void_p),
        ("constants", ctypes.c_void_p), ("volumes", ctypes.c_void_p),
        ("seeds", ctypes.c_void_p), ("n", ctypes.c_uint64), ("zlev", ctypes.c_void_p),
    ]


def build() -> str:
    """Compiles liboracle.so with the committed Makefile (g++, no GPU)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_run.restype = ctypes.c_void_p
        L.oracle_run.argtypes = [ctypes.POINTER(_Args), ctypes.c_char_p, ctypes.c_int]
        L.oracle_counts.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_uint64)] * 3
        L.oracle_get.restype = ctypes.c_int64
        L.oracle_get.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]
        L.oracle_free.argtypes = [ctypes.c_void_p]
        L.oracle_eval_expr.restype = ctypes.c_int
        L.oracle_eval_expr.argtypes = [
            ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_char_p),
            ctypes.POINTER(ctypes.c_int32), ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
            ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double), ctypes.c_char_p,
            ctypes.c_int]
        _lib = L
    return _lib


def expand_scale_offset(grid_scale, offset):
    """GridScale / Offset expansion of source/granite/TrajectoryIntegrator.tpp:521-574.

    GridScale with 1/2/3 entries -> (X, Y, Z) with a 2-entry list reusing [0] for Z. Offset
    follows the reference's branch structure, which tests ``GridScale.size()`` (not
    ``Offset.size()``) for the 2- and 3-entry cases.
    """
    gs = [float(x) for x in grid_scale]
    of = [float(x) for x in offset]
    if len(gs) == 1:
        g = [gs[0]] * 3
    elif len(gs) == 2:
        g = [gs[0], gs[1], gs[0]]
    elif len(gs) == 3:
        g = gs[:3]
    else:
        raise RuntimeError("Invalid GridScale dimension.")
    if len(of) == 1:
        o = [of[0]] * 3
    elif len(gs) == 2:
        o = [of[0], of[1], of[0]]
    elif len(gs) == 3:
        if len(of) < 3:
            raise RuntimeError("Invalid Offset dimension.")
        o = of[:3]
    else:
        raise RuntimeError("Invalid Offset dimension.")
    return g, o


def _enum(v, table):
    if isinstance(v, str):
        return table[v]
    if hasattr(v, "value"):
        return int(v.value)
    return int(v)


def run(settings: dict, wind: np.ndarray, seeds: np.ndarray, *, real64: bool = True,
        weights: str = "exact", time64: bool = False, kahan: bool = False,
        topography=None, uplift=None, constants: dict | None = None,
        volumes: dict | None = None, allow_abs: bool = False, nthreads: int = 0,
        zlevels=None) -> dict:
    """Integrates on the CPU.

    wind: float32 array (S, D, ny, nx) for 2D or (S, D, nz, ny, nx) for 3D -- the slices the
    loader returns in order. seeds: (N, D) float32 start positions. Returns a dict of numpy
    arrays shaped like the reference's TrajectorySet exports.
    """
    L = lib()
    wind = np.ascontiguousarray(wind, dtype=np.float32)
    seeds = np.ascontiguousarray(seeds, dtype=np.float32)
    dim = 3 if _enum(settings.get("Space", 0), SPACE) == 0 else 2
    if wind.ndim != dim + 2 or wind.shape[1] != dim:
        raise ValueError(f"wind must be (S, {dim}, ...) for {dim}D, got {wind.shape}")
    if seeds.ndim != 2 or seeds.shape[1] != dim:
        raise ValueError("seeds must be (N, D)")
    n = seeds.shape[0]
    gs, off = expand_scale_offset(settings.get("GridScale", [25, 25, 25]),
                                  settings.get("Offset", [0, 0, 0]))
    a = _Args()
    a.real64 = 1 if real64 else 0
    a.dim = dim
    a.integrator = _enum(settings.get("Integrator", 1), INTEGRATOR)
    a.border = _enum(settings.get("BorderMode", 0), BORDER)
    a.curvature = _enum(settings.get("CurvatureMode", 0), CURVATURE)
    a.abort_mode = _enum(settings.get("AbortMode", 0), ABORT_MODE)
    a.uplift_mode = _enum(settings.get("UpLiftMode", 0), MODE3)
    a.windfield_mode = _enum(settings.get("WindfieldMode", 1), WINDFIELD)
    a.constants_mode = _enum(settings.get("ConstantsMode", 0), MODE3)
    a.volume_mode = _enum(settings.get("AdditionalVolumeMode", 0), MODE3)
    a.reverse = 1 if settings.get("Reverse", False) else 0
    a.weight_mode = WEIGHTS[weights]
    a.time64 = 1 if time64 else 0
    a.kahan = 1 if kahan else 0
    a.sphere_radius = float(settings.get("SphereRadius", 0.0))
    a.dt = float(settings.get("DeltaT", 1.0))
    a.dtd = float(settings.get("DataTimeDistance", 60.0))
    a.max_length = float(settings.get("MaximumLength", 0.0))
    a.max_sim_time = float(settings.get("MaximumSimulationTime", 100.0))
    for d in range(3):
        a.grid_scale[d] = gs[d]
        a.offset[d] = off[d]
    a.save_interval = int(settings.get("SaveInterval", 1))
    a.min_alive = int(settings.get("MinimumAliveParticles", 0))
    a.nslices = wind.shape[0]
    if dim == 3:
        a.nz, a.ny, a.nx = wind.shape[2], wind.shape[3], wind.shape[4]
    else:
        a.nz, a.ny, a.nx = 1, wind.shape[2], wind.shape[3]
    exprs = list(settings.get("AdditionalCompute", []))
    if len(exprs) > 4:
        raise RuntimeError("at most 4 AdditionalCompute programs are supported")
    keep = []
    a.nexpr = len(exprs)
    for i, e in enumerate(exprs):
        b = e.encode()
        keep.append(b)
        a.exprs[i] = b
    a.wind = wind.ctypes.data
    a.seeds = seeds.ctypes.data
    a.n = n
    a.allow_abs = 1 if allow_abs else 0
    a.nthreads = nthreads if nthreads > 0 else (os.cpu_count() or 1)
    if zlevels is not None:
        zlevels = np.ascontiguousarray(zlevels, dtype=np.float32)
        if zlevels.size != a.nz:
            raise ValueError("zlevels needs one height per z node")
        keep.append(zlevels)
        a.zlev = zlevels.ctypes.data
    if topography is not None:
        topography = np.ascontiguousarray(topography, dtype=np.float32)
        keep.append(topography)
        a.topo = topography.ctypes.data
    if uplift is not None:
        uplift = np.ascontiguousarray(uplift, dtype=np.float32)  # (S, N, D)
        keep.append(uplift)
        a.uplift = uplift.ctypes.data
    if constants:
        names = list(constants.keys())[:4]
        arr = np.ascontiguousarray(np.stack([np.asarray(constants[k], np.float32)
                                             for k in names], axis=-2))  # (S, C, N)
        keep.append(arr)
        a.constants = arr.ctypes.data
        a.nconst = len(names)
        for i, k in enumerate(names):
            b = k.encode()
            keep.append(b)
            a.const_names[i] = b
    if volumes:
        names = list(volumes.keys())[:8]
        arr = np.ascontiguousarray(np.stack([np.asarray(volumes[k], np.float32)
                                             for k in names], axis=1))  # (S, V, ...)
        keep.append(arr)
        a.volumes = arr.ctypes.data
        a.nvol = len(names)
        for i, k in enumerate(names):
            b = k.encode()
            keep.append(b)
            a.vol_names[i] = b
    err = ctypes.create_string_buffer(1024)
    h = L.oracle_run(ctypes.byref(a), err, 1024)
    if not h:
        raise RuntimeError(err.value.decode())
    try:
        steps, cols, ccols = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        L.oracle_counts(h, ctypes.byref(steps), ctypes.byref(cols), ctypes.byref(ccols))

        def get(which, dtype):
            cnt = L.oracle_get(h, which, None)
            out = np.empty(cnt, dtype=dtype)
            if cnt:
                L.oracle_get(h, which, out.ctypes.data)
            return out

        T = cols.value
        pos = get(0, np.float64).reshape(T, n, dim)
        res = {
            "steps": steps.value,
            "positions": pos,                                   # (T, N, D), col 0 = start
            "trajectories": np.ascontiguousarray(pos.transpose(1, 0, 2)),  # (N, T, D)
            "status": get(4, np.int32),
            "abort_step": get(5, np.int64),
        }
        curv = get(1, np.float64)
        res["curvatures"] = curv.reshape(ccols.value, n).T.copy() if ccols.value else \
            np.zeros((n, 0))
        ex = get(2, np.float64)
        ne = len(exprs)
        if ne:
            ex = ex.reshape(T - 1, ne, n)
            res["compute"] = [ex[:, e, :].T.copy() for e in range(ne)]  # (N, cols)
        else:
            res["compute"] = []
        vo = get(3, np.float64)
        nv = int(a.nvol)
        if nv and vo.size:
            vo = vo.reshape(T - 1, nv, n)
            res["volumes"] = {k: vo[:, i, :].T.copy()
                              for i, k in enumerate(list(volumes.keys())[:nv])}
        else:
            res["volumes"] = {}
        res["total_curvature"] = get(6, np.float64)
        res["lengths"] = get(7, np.float64)
        return res
    finally:
        L.oracle_free(h)


def eval_expr(src: str, ref, env: dict | None = None, allow_abs: bool = False):
    """Parses and evaluates one expression on a 19-slot reference vector.

    Returns (float32 result, float64 result); raises RuntimeError on parse errors.
    env maps identifiers to node numbers (volumes: 5 + k, constants: 13 + k).
    """
    L = lib()
    env = env or {}
    names = (ctypes.c_char_p * max(1, len(env)))(*[k.encode() for k in env])
    types = (ctypes.c_int32 * max(1, len(env)))(*[int(v) for v in env.values()])
    r = (ctypes.c_double * 19)(*[float(x) for x in ref])
    o32, o64 = ctypes.c_float(), ctypes.c_double()
    err = ctypes.create_string_buffer(512)
    rc = L.oracle_eval_expr(src.encode(), len(env), names, types, 1 if allow_abs else 0, r,
                            ctypes.byref(o32), ctypes.byref(o64), err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return o32.value, o64.value
