"""Object lifetimes through the C ABI: a context with live fields or jobs is not destroyed, and
the Python Context closes its children before itself."""
import numpy as np
import pytest

from pygranite_amd import capi

pytestmark = pytest.mark.gpu


def test_context_outlives_children(gpu_available):
    ctx = capi.Context(0)
    f = capi.Field(ctx, 8, 8, 8, 3, capi.F32, nslots=2)
    d = capi.JobDesc()
    d.dim, d.integrator, d.dt, d.dtd = 3, 1, 0.1, 1.0
    for k in range(3):
        d.grid_scale[k], d.domain[k] = 1.0, 8.0
    j = capi.Job(ctx, d, np.ones((4, 3), np.float32))
    # the raw destroy refuses while children live
    assert capi.lib().gr_ctx_destroy(ctx.h) == -5
    assert b"still alive" in capi.lib().gr_last_error()
    # closing the wrapper closes the job and the field first
    ctx.close()
    assert not j.h and not f.h and not ctx.h
    # their own close / __del__ are now no-ops
    j.close()
    f.close()


def test_integrators_on_two_threads(pg, gpu_available):
    """Two Python threads run TrajectoryIntegrator.compute() at the same time (each integrator
    owns its context; the GIL is released around device work): each result equals the same
    integration run alone."""
    import math
    import threading

    from helpers import run_product
    from pygranite_amd import synthetic

    wind, gs = synthetic.abc_flow(24, 4)
    st = dict(Space="Space3D", Integrator="ClassicRungeKutta", DeltaT=0.01, DataTimeDistance=1.0,
              GridScale=gs, Offset=[0.0, 0.0, 0.0], MaximumSimulationTime=2.5, SaveInterval=5,
              CurvatureMode="IndividualAndTotalCurvature", AdditionalCompute=["sqrt(u*u+v*v+w*w)"])
    seeds = [synthetic.uniform_seeds(60 + k, 6000, [0.0] * 3, [2 * math.pi] * 3) for k in range(2)]
    alone = [run_product(pg, st, wind, s) for s in seeds]
    got, errs = [None, None], []

    def work(k):
        try:
            for _ in range(3):
                got[k] = run_product(pg, st, wind, seeds[k])
        except Exception as e:  # reported below
            errs.append(e)

    th = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errs, errs
    for k in range(2):
        for f in ("trajectories", "status", "abort_step", "curvatures", "total_curvature"):
            np.testing.assert_array_equal(got[k][f], alone[k][f], err_msg=f)
        np.testing.assert_array_equal(got[k]["compute"][0], alone[k]["compute"][0])


def test_copy_bandwidth_measures_the_device(gpu_available):
    """The streaming-copy measurement returns a device-to-device rate between 1 TB/s and the
    published 8 TB/s peak plus margin, and refuses sizes that are not whole 16-byte units."""
    from pygranite_amd import capi
    ctx = capi.Context(0)
    gbs = ctx.copy_bandwidth(1 << 28, 3)
    assert 1000.0 < gbs < 10000.0, gbs
    import pytest
    with pytest.raises(capi.GraniteError):
        ctx.copy_bandwidth(100, 1)
    ctx.close()


def test_released_job_memory_is_reused_with_identical_results(gpu_available):
    """A destroyed job's arrays (state, ring, index map, shadows) go to the context's cache and
    the next job of the same sizes takes them: its results equal those of the same job on fresh
    memory (cache off), whatever the reused arrays held; the cache keeps at most its cap."""
    import math
    from pygranite_amd import synthetic

    src, gs = synthetic.abc_flow(24, 4)
    seeds = synthetic.uniform_seeds(3, 50000, [0.0] * 3, [2 * math.pi] * 3)

    def desc():
        d = capi.JobDesc()
        d.dim, d.integrator, d.curvature, d.windfield_mode, d.dt, d.dtd = 3, 1, 3, 1, 0.01, 1.0
        for k in range(3):
            d.grid_scale[k] = gs[k]
            d.domain[k] = float(np.float32(24) * np.float32(gs[k]))
        d.save_interval = 1
        d.nprog = 1
        d.prog[0] = capi.compile_expr("sqrt(u*u+v*v+w*w)", dim=3, curvature=3)
        return d

    def run(ctx, f):
        job = capi.Job(ctx, desc(), seeds, save_capacity=60)
        job.reorder()
        out, t, counter = [], np.float32(0), 0
        for steps in (60, 40):  # two launches in window 0 (slices 0 and 1)
            w = capi.Window()
            w.wind, w.wind_back, w.wind_front = f.h, 0, 1
            w.time_back, w.time_front = 0.0, 1.0
            w.time_back64, w.time_front64 = 0.0, 1.0
            w.time_start = w.time_start64 = float(t)
            w.counter_start, w.nsteps = counter, steps
            job.advance(w)
            out += [job.read_saved(capi.OUT_POSITIONS, counter, steps), job.read_saved(capi.OUT_EXPR0, counter, steps)]
            for _ in range(steps):
                t = np.float32(t + np.float32(0.01))
            counter += steps
            job.reorder()
        out += [job.read(capi.OUT_POSITIONS), job.read(capi.OUT_STATUS), job.read(capi.OUT_ABORT_STEP),
                job.read(capi.OUT_TOTAL_CURVATURE)]
        job.close()
        return out

    outs = {}
    for cap in (0, 8 << 30):
        ctx = capi.Context(0)
        ctx.set_option(capi.CTX_POOL_BYTES, cap)
        assert ctx.get_option(capi.CTX_POOL_BYTES) == cap
        f = capi.Field(ctx, 24, 24, 24, 3, capi.F32, nslots=3)
        for sl in range(3):
            f.upload(sl, [src[sl, c] for c in range(3)])
        first = run(ctx, f)
        held = ctx.get_option(capi.CTX_POOL_CACHED)
        second = run(ctx, f)  # on the first job's memory when the cache is on
        if cap:
            assert held >= 50000 * 4 * 20  # the first job's arrays, shadows and rings
            assert ctx.get_option(capi.CTX_POOL_CACHED) == held  # taken and given back
        else:
            assert held == 0 and ctx.get_option(capi.CTX_POOL_CACHED) == 0
        for a, b in zip(first, second):
            np.testing.assert_array_equal(a, b)
        outs[cap] = first
        with pytest.raises(capi.GraniteError):
            ctx.set_option(capi.CTX_POOL_CACHED, 0)
        ctx.set_option(capi.CTX_POOL_BYTES, 0)  # a cap below the cached bytes empties the cache
        assert ctx.get_option(capi.CTX_POOL_CACHED) == 0
        f.close()
        ctx.close()
    for a, b in zip(outs[0], outs[8 << 30]):
        np.testing.assert_array_equal(a, b)
    assert (outs[0][-3] != 0).sum() > 1000  # particles stopped on the way


_TORCH_SEEDS = r"""
import math, sys
import numpy as np
import torch
torch.zeros(1, device="cuda")  # torch's runtime serves the process, as in bench.py
sys.path.insert(0, sys.argv[1])
from pygranite_amd import capi, synthetic

src, gs = synthetic.abc_flow(24, 3)
ctx = capi.Context(0)
f = capi.Field(ctx, 24, 24, 24, 3, capi.F32, nslots=2)
for sl in range(2):
    f.upload(sl, [src[sl, c] for c in range(3)])
d = capi.JobDesc()
d.dim, d.integrator, d.curvature, d.windfield_mode, d.dt, d.dtd = 3, 1, 0, 1, 0.01, 1.0
for k in range(3):
    d.grid_scale[k] = gs[k]
    d.domain[k] = float(np.float32(24) * np.float32(gs[k]))
n = 1 << 20
g = torch.Generator(device="cuda")
g.manual_seed(7)
t = torch.rand((n, 3), generator=g, device="cuda", dtype=torch.float64)
for _ in range(8):  # a queue of torch kernels ahead of the seeds' last write
    t = torch.sin(t) * 0.5 + 0.5
host = (t * (2 * math.pi)).float().cpu().numpy()
seeds_t = (t * (2 * math.pi)).float().contiguous()  # written again, still queued
a = capi.Job(ctx, d, seeds_t)
del seeds_t, t
junk = [torch.full((n, 3), -1e30, device="cuda") for _ in range(4)]  # reuse freed blocks
b = capi.Job(ctx, d, host)
for job in (a, b):
    w = capi.Window()
    w.wind, w.wind_back, w.wind_front = f.h, 0, 1
    w.time_back, w.time_front = 0.0, 1.0
    w.time_back64, w.time_front64 = 0.0, 1.0
    w.counter_start, w.nsteps = 0, 20
    job.advance(w)
pa, pb = a.read(capi.OUT_POSITIONS), b.read(capi.OUT_POSITIONS)
sa, sb = a.read(capi.OUT_STATUS), b.read(capi.OUT_STATUS)
print("equal", bool(np.array_equal(pa, pb) and np.array_equal(sa, sb)), int((sb != 0).sum()), flush=True)
del junk
a.close()
b.close()
f.close()
ctx.close()
"""


def test_device_seeds_from_torch_equal_host_seeds(gpu_available):
    """Seeds produced by torch kernels just before gr_job_create, and the tensor dropped and its
    memory overwritten right after, give the same job as the same seeds from host memory: the
    binding waits for torch's stream and keeps the tensor until the job is closed. Runs in a
    process of its own whose HIP runtime torch loads first (as bench.py does)."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", _TORCH_SEEDS, root], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-3000:]
    line = [l for l in r.stdout.splitlines() if l.startswith("equal")][-1].split()
    assert line[1] == "True", r.stdout


def test_context_option_validation(gpu_available):
    """Context options reject values outside their ranges and the read-only ones reject writes;
    defaults as documented."""
    ctx = capi.Context(0)
    assert ctx.get_option(capi.CTX_BATCH_UNIFORM) == 1
    assert ctx.get_option(capi.CTX_POOL_BYTES) == 8 << 30
    assert ctx.get_option(capi.CTX_POOL_CACHED) == 0
    assert ctx.get_option(capi.CTX_BATCH_LAST_KERNEL) == 0
    assert ctx.get_option(capi.CTX_BATCH_QUAD_ITEMS) == 0
    for opt, bad in ((capi.CTX_BATCH_UNIFORM, 2), (capi.CTX_POOL_BYTES, -1), (capi.CTX_BATCH_LAST_KERNEL, 0),
                     (capi.CTX_BATCH_QUAD_ITEMS, 0), (capi.CTX_POOL_CACHED, 0), (99, 0)):
        with pytest.raises(capi.GraniteError):
            ctx.set_option(opt, bad)
    ctx.set_option(capi.CTX_BATCH_UNIFORM, 0)
    assert ctx.get_option(capi.CTX_BATCH_UNIFORM) == 0
    with pytest.raises(capi.GraniteError):
        ctx.get_option(99)
    ctx.close()
