"""ctypes binding of the C ABI in include/granite_hip.h (libgranite_hip.so).

This is the layer tests and the benchmark drive directly: it exposes fields (resident slices),
jobs (resident particle state) and windows exactly as the C ABI defines them, without the
pygranite host schedule on top. Structures mirror the header field by field.
"""
from __future__ import annotations

import ctypes
import weakref
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "libgranite_hip.so")

MAX_CODE = 64
MAX_PROGRAMS = 4
F32, F16, BF16 = 0, 1, 2
OUT_POSITIONS, OUT_STATUS, OUT_TOTAL_CURVATURE, OUT_LENGTHS = 0, 1, 2, 3
OUT_CURVATURE, OUT_ABORT_STEP, OUT_EXPR0, OUT_VOLUME0 = 4, 5, 8, 16
EXPR_ALLOW_ABS = 1

# every symbol the header declares (checked by the CPU test suite)
EXPORTS = [
    "gr_abi_version", "gr_last_error", "gr_device_count", "gr_ctx_create", "gr_ctx_destroy",
    "gr_ctx_synchronize", "gr_ctx_info", "gr_ctx_take_work", "gr_ctx_take_tile_stats", "gr_ctx_copy_bandwidth", "gr_ctx_take_kernel_ms", "gr_ctx_set_option",
    "gr_ctx_get_option", "gr_field_create", "gr_field_destroy",
    "gr_field_upload", "gr_field_upload_device", "gr_field_wait", "gr_field_bytes",
    "gr_field_set_z_levels",
    "gr_expr_compile", "gr_expr_eval_host", "gr_expr_eval_device_form", "gr_job_create", "gr_job_destroy", "gr_job_bytes",
    "gr_job_set_particle_data", "gr_job_advance", "gr_job_advance_noalloc", "gr_job_read", "gr_job_read_saved",
    "gr_job_read_saved_async", "gr_job_read_wait",
    "gr_job_count_alive", "gr_job_live_range", "gr_job_saved_columns", "gr_job_last_advance_ms", "gr_job_set_launch",
    "gr_job_set_option", "gr_job_reorder", "gr_jobs_reorder", "gr_batch_run", "gr_batch_autotune",
    "gr_job_autotune",
]
OPT_BLOCK, OPT_PAIRED, OPT_CU_WAVES, OPT_SPECIALIZE, OPT_REUSE, OPT_QUAD = 1, 2, 3, 4, 5, 6
OPT_TILE, OPT_TILE_HALO = 7, 8
CTX_BATCH_BLOCK, CTX_BATCH_WAVES, CTX_BATCH_GROUP, CTX_BATCH_CU_WAVES, CTX_BATCH_REUSE = 1, 2, 3, 4, 5
CTX_BATCH_QUAD_ITEMS = 6  # read-only: quad items of the last batched launch
CTX_BATCH_UNIFORM = 7  # 1: per-variant kernels for stencil-reuse launches of one variant
CTX_BATCH_LAST_KERNEL = 8  # read-only: 0 none, 1 default, 2 128-VGPR, 3 stencil-reuse, 4 per-variant
CTX_POOL_BYTES, CTX_POOL_CACHED = 9, 10  # cache of destroyed jobs' arrays: cap, bytes held (read-only)


class FieldDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_uint32), ("ny", ctypes.c_uint32), ("nz", ctypes.c_uint32),
                ("ncomp", ctypes.c_uint32), ("dtype", ctypes.c_uint32),
                ("nslots", ctypes.c_uint32), ("address", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class Program(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("depth", ctypes.c_uint32),
                ("code", ctypes.c_uint32 * MAX_CODE), ("imm", ctypes.c_float * MAX_CODE)]


class ExprEnv(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_uint32), ("curvature", ctypes.c_uint32),
                ("nconst", ctypes.c_uint32), ("nvol", ctypes.c_uint32),
                ("const_names", ctypes.POINTER(ctypes.c_char_p)),
                ("vol_names", ctypes.POINTER(ctypes.c_char_p)), ("flags", ctypes.c_uint32)]


class JobDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_uint32), ("integrator", ctypes.c_uint32),
                ("border", ctypes.c_uint32), ("curvature", ctypes.c_uint32),
                ("abort_mode", ctypes.c_uint32), ("uplift_mode", ctypes.c_uint32),
                ("windfield_mode", ctypes.c_uint32), ("constants_mode", ctypes.c_uint32),
                ("volume_mode", ctypes.c_uint32), ("reverse", ctypes.c_uint32),
                ("weights", ctypes.c_uint32), ("time64", ctypes.c_uint32),
                ("kahan", ctypes.c_uint32), ("local_coords", ctypes.c_uint32),
                ("nvol", ctypes.c_uint32), ("nconst", ctypes.c_uint32),
                ("dt", ctypes.c_float), ("dtd", ctypes.c_float),
                ("sphere_radius", ctypes.c_float), ("max_length", ctypes.c_float),
                ("grid_scale", ctypes.c_float * 3), ("domain", ctypes.c_float * 3),
                ("offset", ctypes.c_double * 3), ("save_interval", ctypes.c_uint64),
                ("nprog", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("prog", Program * MAX_PROGRAMS)]


class Window(ctypes.Structure):
    _fields_ = [("wind", ctypes.c_void_p), ("wind_back", ctypes.c_uint32),
                ("wind_front", ctypes.c_uint32), ("volumes", ctypes.c_void_p),
                ("vol_back", ctypes.c_uint32), ("vol_front", ctypes.c_uint32),
                ("topography", ctypes.c_void_p), ("uplift_back", ctypes.c_uint32),
                ("uplift_front", ctypes.c_uint32), ("const_back", ctypes.c_uint32),
                ("const_front", ctypes.c_uint32), ("time_back", ctypes.c_float),
                ("time_front", ctypes.c_float), ("time_back64", ctypes.c_double),
                ("time_front64", ctypes.c_double), ("time_start", ctypes.c_float),
                ("time_start64", ctypes.c_double), ("counter_start", ctypes.c_uint64),
                ("nsteps", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class Span(ctypes.Structure):
    _fields_ = [("job", ctypes.c_void_p), ("wind", ctypes.c_void_p),
                ("first_window", ctypes.c_uint32), ("max_steps", ctypes.c_uint32),
                ("time_start", ctypes.c_float), ("max_sim_time", ctypes.c_float),
                ("time_start64", ctypes.c_double), ("counter_start", ctypes.c_uint64)]


class SpanEnd(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_uint32), ("window", ctypes.c_uint32),
                ("time", ctypes.c_float), ("done", ctypes.c_uint32),
                ("time64", ctypes.c_double), ("counter", ctypes.c_uint64)]


class TuneEntry(ctypes.Structure):
    _fields_ = [("block", ctypes.c_uint32), ("waves", ctypes.c_uint32), ("group", ctypes.c_uint32),
                ("ms", ctypes.c_float), ("cu_waves", ctypes.c_uint32)]


class JobTuneEntry(ctypes.Structure):
    _fields_ = [("block", ctypes.c_uint32), ("lanes", ctypes.c_uint32), ("tile_bytes", ctypes.c_uint32),
                ("ms", ctypes.c_float)]


_lib = None


def lib():
    """Loads libgranite_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBRARY_PATH):
        raise ImportError(f"{LIBRARY_PATH} missing: run `python -m pygranite_amd.build`")
    L = ctypes.CDLL(LIBRARY_PATH)
    vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    P = ctypes.POINTER
    sig = {
        "gr_abi_version": ([], i32),
        "gr_last_error": ([], ctypes.c_char_p),
        "gr_device_count": ([P(i32)], i32),
        "gr_ctx_create": ([i32, P(vp)], i32),
        "gr_ctx_destroy": ([vp], i32),
        "gr_ctx_synchronize": ([vp], i32),
        "gr_ctx_info": ([vp, ctypes.c_char_p, ctypes.c_size_t, P(u64), P(i32)], i32),
        "gr_ctx_take_work": ([vp, P(u64)], i32),
        "gr_ctx_take_tile_stats": ([vp, P(u64), P(u64)], i32),
        "gr_ctx_copy_bandwidth": ([vp, u64, u32, P(ctypes.c_double)], i32),
        "gr_ctx_take_kernel_ms": ([vp, P(ctypes.c_double), P(u64)], i32),
        "gr_ctx_set_option": ([vp, u32, ctypes.c_int64], i32),
        "gr_ctx_get_option": ([vp, u32, P(ctypes.c_int64)], i32),
        "gr_field_create": ([vp, P(FieldDesc), P(vp)], i32),
        "gr_field_destroy": ([vp], i32),
        "gr_field_upload": ([vp, u32, P(vp), i32], i32),
        "gr_field_upload_device": ([vp, u32, P(vp)], i32),
        "gr_field_wait": ([vp, u32], i32),
        "gr_field_bytes": ([vp, P(u64)], i32),
        "gr_field_set_z_levels": ([vp, P(ctypes.c_float), u32], i32),
        "gr_expr_compile": ([ctypes.c_char_p, P(ExprEnv), P(Program)], i32),
        "gr_expr_eval_host": ([P(Program), P(ctypes.c_float), P(ctypes.c_float)], i32),
        "gr_expr_eval_device_form": ([P(Program), P(ctypes.c_float), P(ctypes.c_float)], i32),
        "gr_job_create": ([vp, P(JobDesc), vp, u64, u32, P(vp)], i32),
        "gr_job_destroy": ([vp], i32),
        "gr_job_bytes": ([P(JobDesc), u64, u32, i32, P(u64)], i32),
        "gr_job_set_particle_data": ([vp, u32, u32, vp], i32),
        "gr_job_advance": ([vp, P(Window)], i32),
        "gr_job_advance_noalloc": ([vp, P(Window)], i32),
        "gr_job_read": ([vp, u32, vp, u64], i32),
        "gr_job_read_saved": ([vp, u32, u64, u64, vp], i32),
        "gr_job_read_saved_async": ([vp, u32, u64, u64, P(vp), P(u64)], i32),
        "gr_job_read_wait": ([vp, u64], i32),
        "gr_job_count_alive": ([vp, P(u64)], i32),
        "gr_job_live_range": ([vp, P(u64)], i32),
        "gr_job_saved_columns": ([vp, P(u64)], i32),
        "gr_job_last_advance_ms": ([vp, P(ctypes.c_float)], i32),
        "gr_job_set_launch": ([vp, u32], i32),
        "gr_job_set_option": ([vp, u32, ctypes.c_int64], i32),
        "gr_job_reorder": ([vp], i32),
        "gr_jobs_reorder": ([P(vp), u32], i32),
        "gr_batch_run": ([vp, P(Span), u32, u32, P(SpanEnd), P(ctypes.c_float)], i32),
        "gr_batch_autotune": ([vp, vp, P(JobDesc), u32, u32, u32, u32, P(TuneEntry), u32, P(u32)], i32),
        "gr_job_autotune": ([vp, vp, u32, u32, P(JobTuneEntry), u32, P(u32)], i32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


class GraniteError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise GraniteError(f"granite error {rc}: {lib().gr_last_error().decode()}")


def job_bytes(desc, n: int, save_capacity: int = 1, reorder: bool = True) -> int:
    """Device bytes of a job (gr_job_bytes): what gr_job_create allocates, plus what its first
    reorder adds when `reorder`."""
    b = ctypes.c_uint64()
    _check(lib().gr_job_bytes(ctypes.byref(desc), n, save_capacity, 1 if reorder else 0, ctypes.byref(b)))
    return b.value


def device_count() -> int:
    n = ctypes.c_int(0)
    _check(lib().gr_device_count(ctypes.byref(n)))
    return n.value


def compile_expr(src: str, dim: int = 3, curvature: int = 0, constants=(), volumes=(),
                 allow_abs: bool = False) -> Program:
    env = ExprEnv()
    env.dim = dim
    env.curvature = curvature
    cn = (ctypes.c_char_p * max(1, len(constants)))(*[c.encode() for c in constants])
    vn = (ctypes.c_char_p * max(1, len(volumes)))(*[v.encode() for v in volumes])
    env.nconst = len(constants)
    env.nvol = len(volumes)
    env.const_names = cn
    env.vol_names = vn
    env.flags = EXPR_ALLOW_ABS if allow_abs else 0
    prog = Program()
    rc = lib().gr_expr_compile(src.encode(), ctypes.byref(env), ctypes.byref(prog))
    if rc != 0:
        raise GraniteError(lib().gr_last_error().decode())
    return prog


def eval_expr_host(prog: Program, ref) -> float:
    r = (ctypes.c_float * 19)(*[float(x) for x in ref])
    out = ctypes.c_float()
    _check(lib().gr_expr_eval_host(ctypes.byref(prog), r, ctypes.byref(out)))
    return out.value


def eval_expr_device_form(prog: Program, ref) -> float:
    """The program evaluated in the device form the step kernels run (host arithmetic)."""
    r = (ctypes.c_float * 19)(*[float(x) for x in ref])
    out = ctypes.c_float()
    _check(lib().gr_expr_eval_device_form(ctypes.byref(prog), r, ctypes.byref(out)))
    return out.value


class Context:
    """A device context. Fields and jobs made from it hold a reference to it; closing the
    context closes them first (the library refuses to destroy a context with live children)."""

    def __init__(self, device: int = 0):
        self.h = ctypes.c_void_p()
        self._children = weakref.WeakSet()
        _check(lib().gr_ctx_create(device, ctypes.byref(self.h)))

    def _adopt(self, child):
        self._children.add(child)

    def batch_run(self, spans, block: int = 0, timed: bool = True, ends: bool = False):
        """Runs [(job, field, dict(first_window=, max_steps=, time_start=, time_start64=,
        max_sim_time=, counter_start=))...] in one persistent launch.

        Returns (steps per span, device ms), or (list of end records, ms) with ends=True."""
        arr = (Span * max(1, len(spans)))()
        for i, (job, field, kw) in enumerate(spans):
            s = arr[i]
            s.job, s.wind = job.h, field.h
            s.first_window = kw.get("first_window", 0)
            s.max_steps = kw["max_steps"]
            s.time_start = kw.get("time_start", 0.0)
            s.time_start64 = kw.get("time_start64", kw.get("time_start", 0.0))
            s.max_sim_time = kw.get("max_sim_time", 0.0)
            s.counter_start = kw.get("counter_start", 0)
        out = (SpanEnd * max(1, len(spans)))()
        ms = ctypes.c_float()
        _check(lib().gr_batch_run(self.h, arr, len(spans), block, out,
                                  ctypes.byref(ms) if timed else None))
        recs = [dict(steps=e.steps, window=e.window, time=e.time, time64=e.time64,
                     counter=e.counter, done=bool(e.done)) for e in list(out)[:len(spans)]]
        if ends:
            return recs, ms.value
        return [r["steps"] for r in recs], ms.value

    def info(self):
        name = ctypes.create_string_buffer(256)
        hbm = ctypes.c_uint64()
        cus = ctypes.c_int()
        _check(lib().gr_ctx_info(self.h, name, 256, ctypes.byref(hbm), ctypes.byref(cus)))
        return {"name": name.value.decode(), "hbm_bytes": hbm.value, "compute_units": cus.value}

    def synchronize(self):
        _check(lib().gr_ctx_synchronize(self.h))

    def set_option(self, option: int, value: int):
        _check(lib().gr_ctx_set_option(self.h, option, value))

    def get_option(self, option: int) -> int:
        v = ctypes.c_int64(0)
        _check(lib().gr_ctx_get_option(self.h, option, ctypes.byref(v)))
        return int(v.value)

    def autotune(self, field, desc, n: int = 65536, njobs: int = 1, steps: int = 50, trials: int = 2):
        """Runs the batched-kernel autotuner on njobs probe jobs of n particles of `desc` over
        `field`; sets the fastest configuration on this context and returns every candidate's
        device time."""
        res = (TuneEntry * 64)()
        cnt = ctypes.c_uint32(0)
        _check(lib().gr_batch_autotune(self.h, field.h, ctypes.byref(desc), n, njobs, steps, trials,
                                       res, 64, ctypes.byref(cnt)))
        return [{"block": r.block, "waves": r.waves, "group": r.group, "cu_waves": r.cu_waves,
                 "ms": r.ms} for r in res[:cnt.value]]

    def reorder(self, jobs):
        """Reorders several jobs of this context in one pass (gr_jobs_reorder)."""
        arr = (ctypes.c_void_p * max(1, len(jobs)))(*[j.h for j in jobs])
        _check(lib().gr_jobs_reorder(arr, len(jobs)))

    def take_kernel_ms(self):
        """(summed device ms, launches) of the step kernels since the previous call (then reset)."""
        ms, n = ctypes.c_double(0), ctypes.c_uint64(0)
        _check(lib().gr_ctx_take_kernel_ms(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return float(ms.value), int(n.value)

    def copy_bandwidth(self, nbytes: int = 1 << 30, trials: int = 5) -> float:
        """Device-to-device copy bandwidth in GB/s (read + write), best of `trials`."""
        v = ctypes.c_double()
        _check(lib().gr_ctx_copy_bandwidth(self.h, nbytes, trials, ctypes.byref(v)))
        return v.value

    def take_tile_stats(self):
        """(workgroup-steps with a field tile, without one) since the last call (resets both)."""
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().gr_ctx_take_tile_stats(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def take_work(self) -> int:
        """Integrated particle-steps of every launch since the previous call (then reset)."""
        v = ctypes.c_uint64(0)
        _check(lib().gr_ctx_take_work(self.h, ctypes.byref(v)))
        return int(v.value)

    def close(self):
        if self.h:
            for child in list(self._children):
                child.close()
            _check(lib().gr_ctx_destroy(self.h))
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Field:
    def __init__(self, ctx: Context, nx, ny, nz, ncomp, dtype=F32, nslots=3, address=0):
        self.ctx = ctx
        d = FieldDesc(nx, ny, nz, ncomp, dtype, nslots, address, 0)
        self.desc = d
        self.h = ctypes.c_void_p()
        _check(lib().gr_field_create(ctx.h, ctypes.byref(d), ctypes.byref(self.h)))
        ctx._adopt(self)

    def upload(self, slot: int, comps, async_: bool = False):
        arrs = [np.ascontiguousarray(c, dtype=np.float32) for c in comps]
        ptrs = (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        _check(lib().gr_field_upload(self.h, slot, ptrs, 1 if async_ else 0))
        self._keep = arrs

    def wait(self, slot: int):
        _check(lib().gr_field_wait(self.h, slot))

    def upload_device(self, slot: int, ptrs):
        """Uploads FP32 component arrays already on the device (e.g. torch tensors' data_ptr())."""
        arr = (ctypes.c_void_p * len(ptrs))(*[int(p) for p in ptrs])
        _check(lib().gr_field_upload_device(self.h, slot, arr))

    def set_z_levels(self, levels):
        z = np.ascontiguousarray(levels, dtype=np.float32)
        _check(lib().gr_field_set_z_levels(self.h, z.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                           z.size))

    def bytes_per_slot(self) -> int:
        b = ctypes.c_uint64()
        _check(lib().gr_field_bytes(self.h, ctypes.byref(b)))
        return b.value

    def close(self):
        if self.h:
            lib().gr_field_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Job:
    def __init__(self, ctx: Context, desc: JobDesc, seeds, save_capacity: int = 1):
        """seeds: (n, dim) float32 numpy array, or a contiguous float32 torch tensor on the
        context's GPU (copied on the device)."""
        self.ctx = ctx
        self.desc = desc
        self.dim = desc.dim
        self._seeds = None
        if hasattr(seeds, "data_ptr") and getattr(seeds, "is_cuda", False):
            if str(seeds.dtype) != "torch.float32" or not seeds.is_contiguous():
                raise ValueError("device seeds must be a contiguous float32 tensor")
            import torch
            # the engine copies device seeds on its own stream, ordered after nothing of torch's:
            # the tensor's producer must be done, and the tensor must outlive the queued copy
            torch.cuda.current_stream(seeds.device).synchronize()
            self._seeds = seeds
            ptr, self.n = seeds.data_ptr(), int(seeds.shape[0])
        else:
            seeds = np.ascontiguousarray(seeds, dtype=np.float32)
            ptr, self.n = seeds.ctypes.data, seeds.shape[0]
        self.h = ctypes.c_void_p()
        _check(lib().gr_job_create(ctx.h, ctypes.byref(desc), ptr, self.n,
                                   save_capacity, ctypes.byref(self.h)))
        ctx._adopt(self)

    def set_launch(self, block: int):
        _check(lib().gr_job_set_launch(self.h, block))

    def set_option(self, option: int, value: int):
        _check(lib().gr_job_set_option(self.h, option, value))

    def autotune(self, field, steps: int = 50, trials: int = 2):
        """Runs the single-job autotuner (gr_job_autotune) on a probe job of this job's
        description over `field`; sets the fastest block size, lanes per particle and field tile
        on this job and returns every candidate's device time."""
        res = (JobTuneEntry * 64)()
        cnt = ctypes.c_uint32(0)
        _check(lib().gr_job_autotune(self.h, field.h, steps, trials, res, 64, ctypes.byref(cnt)))
        return [{"block": r.block, "lanes": r.lanes, "tile_bytes": r.tile_bytes, "ms": r.ms}
                for r in res[:cnt.value]]

    def reorder(self):
        _check(lib().gr_job_reorder(self.h))

    def set_particle_data(self, kind: int, slot: int, data: np.ndarray):
        a = np.ascontiguousarray(data, dtype=np.float32)
        _check(lib().gr_job_set_particle_data(self.h, kind, slot, a.ctypes.data))

    def advance(self, w: Window):
        _check(lib().gr_job_advance(self.h, ctypes.byref(w)))

    def advance_noalloc(self, w: Window):
        """gr_job_advance_noalloc: the same window without any allocation."""
        _check(lib().gr_job_advance_noalloc(self.h, ctypes.byref(w)))

    def read(self, kind: int) -> np.ndarray:
        if kind == OUT_POSITIONS:
            out = np.empty((self.n, self.dim), np.float32)
        elif kind in (OUT_STATUS, OUT_ABORT_STEP):
            out = np.empty(self.n, np.int32)
        else:
            out = np.empty(self.n, np.float32)
        _check(lib().gr_job_read(self.h, kind, out.ctypes.data, out.nbytes))
        return out

    def read_saved(self, kind: int, first: int, count: int) -> np.ndarray:
        if kind == OUT_POSITIONS:
            out = np.empty((count, self.n, self.dim), np.float32)
        else:
            out = np.empty((count, self.n), np.float32)
        _check(lib().gr_job_read_saved(self.h, kind, first, count, out.ctypes.data))
        return out

    def read_saved_async(self, kind: int, first: int, count: int):
        """Starts gr_job_read_saved_async into a new array; returns (ticket, array). The array
        holds the columns once wait(ticket) has returned."""
        if kind == OUT_POSITIONS:
            out = np.empty((count, self.n, self.dim), np.float32)
        else:
            out = np.empty((count, self.n), np.float32)
        cols = (ctypes.c_void_p * max(1, count))(*[out[c].ctypes.data for c in range(count)])
        t = ctypes.c_uint64()
        _check(lib().gr_job_read_saved_async(self.h, kind, first, count, cols, ctypes.byref(t)))
        return t.value, out

    def wait(self, ticket: int):
        _check(lib().gr_job_read_wait(self.h, ticket))

    def saved_columns(self) -> int:
        c = ctypes.c_uint64()
        _check(lib().gr_job_saved_columns(self.h, ctypes.byref(c)))
        return c.value

    def count_alive(self) -> int:
        c = ctypes.c_uint64()
        _check(lib().gr_job_count_alive(self.h, ctypes.byref(c)))
        return c.value

    def live_range(self) -> int:
        """Particle slots the next launch covers (every later slot holds a stopped particle)."""
        c = ctypes.c_uint64()
        _check(lib().gr_job_live_range(self.h, ctypes.byref(c)))
        return c.value

    def last_advance_ms(self) -> float:
        ms = ctypes.c_float()
        _check(lib().gr_job_last_advance_ms(self.h, ctypes.byref(ms)))
        return ms.value

    def close(self):
        if self.h:
            lib().gr_job_destroy(self.h)
            self.h = ctypes.c_void_p()
        self._seeds = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
