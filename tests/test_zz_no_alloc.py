"""gr_job_advance_noalloc (the entry point that allocates nothing) against gr_job_advance.

Two calls in a row through the C ABI, the second with a smaller input (fewer particles and
steps, another window): every output -- saved columns, positions, status, abort steps,
curvature, expression -- equals gr_job_advance's bit for bit. The first call finds the window's
time-paired copy resident, the second does not (its slots are sampled separately). Sizes beyond
the scratch array are rejected.
"""
import math

import numpy as np
import pytest

from pygranite_amd import capi, synthetic

pytestmark = pytest.mark.gpu


def _desc(gs, n, kahan=0, time64=0):
    d = capi.JobDesc()
    d.dim, d.integrator, d.curvature, d.windfield_mode = 3, 1, 3, 1
    d.dt, d.dtd = 0.01, 1.0
    d.kahan, d.time64 = kahan, time64
    for k in range(3):
        d.grid_scale[k] = gs[k]
        d.domain[k] = float(np.float32(n) * np.float32(gs[k]))
    d.save_interval = 1
    d.nprog = 1
    d.prog[0] = capi.compile_expr("sqrt(u*u+v*v+w*w)", dim=3, curvature=3)
    return d


def _window(f, c, counter, t, k):
    w = capi.Window()
    w.wind, w.wind_back, w.wind_front = f.h, c, c + 1
    w.time_back, w.time_front = float(c), float(c + 1)
    w.time_back64, w.time_front64 = float(c), float(c + 1)
    w.time_start = w.time_start64 = float(t)
    w.counter_start, w.nsteps = counter, k
    return w


def _outputs(job, k):
    return [job.read_saved(capi.OUT_POSITIONS, 0, k), job.read_saved(capi.OUT_CURVATURE, 0, k),
            job.read_saved(capi.OUT_EXPR0, 0, k), job.read(capi.OUT_POSITIONS), job.read(capi.OUT_STATUS),
            job.read(capi.OUT_ABORT_STEP), job.read(capi.OUT_TOTAL_CURVATURE)]


def test_noalloc_advance_equals_advance_on_two_calls(gpu_available):
    n = 32
    src, gs = synthetic.abc_flow(n, 4)
    ctx = capi.Context(0)
    f = capi.Field(ctx, n, n, n, 3, capi.F16, nslots=4)
    for s in range(4):
        f.upload(s, [src[s, c] for c in range(3)])
    seeds = synthetic.uniform_seeds(41, 3000, [0.0] * 3, [2 * math.pi] * 3)
    # call 1: 3000 particles, 150 steps of window 0 (gr_job_advance first: it builds the pair)
    calls = [(seeds, 0, 0, 0.0, 150, (0, 0), False), (seeds[:700], 1, 101, np.float32(1.01), 60, (1, 1), True)]
    for seeds_i, c, counter, t, k, (kahan, t64), noalloc_first in calls:
        d = _desc(gs, n, kahan, t64)
        a = capi.Job(ctx, d, seeds_i, save_capacity=k)
        b = capi.Job(ctx, d, seeds_i, save_capacity=k)
        w = _window(f, c, counter, t, k)
        if noalloc_first:
            b.advance_noalloc(w)
            a.advance(w)
        else:
            a.advance(w)
            b.advance_noalloc(w)
        ctx.synchronize()
        assert a.saved_columns() == b.saved_columns() == k
        for x, y in zip(_outputs(a, k), _outputs(b, k)):
            np.testing.assert_array_equal(x, y)
        assert (a.read(capi.OUT_STATUS) != 0).sum() > 0 or k < 100
        a.close()
        b.close()
    # a window beyond the scratch array is refused before anything runs
    d = _desc(gs, n)
    j = capi.Job(ctx, d, seeds[:64], save_capacity=1)
    with pytest.raises(capi.GraniteError, match="scratch"):
        j.advance_noalloc(_window(f, 0, 0, 0.0, 16385))
    j.close()
    f.close()
    ctx.close()
