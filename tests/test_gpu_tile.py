"""GPU parity of the field-tile step kernel (GR_OPT_TILE): every workgroup stages the box of
time-paired nodes around its live particles' cells in LDS each step and samples the stencils
inside it from there. The words are the same as the paired copy's, so every output must equal
the untiled time-paired kernel's bit for bit -- which the full-size tests pin to the FP32
oracle -- whatever the tile size, halo, workgroup size and particle order; the tile statistics
show that the tile was used (or, for a tile too small for the boxes, that the fallback was)."""
import math

import numpy as np
import pytest

import oracle
from helpers import assert_bitexact
from pygranite_amd import capi, synthetic

pytestmark = pytest.mark.gpu


def _desc(gs, n, integrator, curv, kahan, expr="sqrt(u*u+v*v+w*w)*2-x", dt=0.01):
    d = capi.JobDesc()
    d.dim, d.integrator, d.curvature, d.windfield_mode = 3, integrator, curv, 1
    d.dt, d.dtd = dt, 1.0
    d.kahan = kahan
    for k in range(3):
        d.grid_scale[k] = gs[k]
        d.domain[k] = float(np.float32(n) * np.float32(gs[k]))
    d.save_interval = 1
    d.nprog = 1
    d.prog[0] = capi.compile_expr(expr, dim=3, curvature=curv)
    return d


def _run(ctx, f, d, seeds, windows, steps, tile, halo=1, block=0, reorder=True, specialize=1):
    """Advances a job through `windows` windows of `steps` steps (regrouped between windows when
    `reorder`), returning every saved column and the final state."""
    job = capi.Job(ctx, d, seeds, save_capacity=steps)
    job.set_option(capi.OPT_QUAD, 0)
    job.set_option(capi.OPT_TILE, tile)
    job.set_option(capi.OPT_TILE_HALO, halo)
    job.set_option(capi.OPT_SPECIALIZE, specialize)
    if block:
        job.set_launch(block)
    if reorder:
        job.reorder()
    counter, t = 0, np.float32(0)
    dt = np.float32(d.dt)
    o = []
    for wb in range(windows):
        w = capi.Window()
        w.wind, w.wind_back, w.wind_front = f.h, wb, wb + 1
        w.time_back, w.time_front = float(wb), float(wb + 1)
        w.time_back64, w.time_front64 = float(wb), float(wb + 1)
        w.time_start = w.time_start64 = float(t)
        w.counter_start, w.nsteps = counter, steps
        job.advance(w)
        o += [job.read_saved(capi.OUT_POSITIONS, counter, steps), job.read_saved(capi.OUT_EXPR0, counter, steps)]
        if d.curvature:
            o.append(job.read_saved(capi.OUT_CURVATURE, counter, steps))
        for _ in range(steps):
            t = np.float32(t + dt)
        counter += steps
        if reorder:
            job.reorder()
    o += [job.read(capi.OUT_STATUS), job.read(capi.OUT_ABORT_STEP), job.read(capi.OUT_POSITIONS)]
    if d.curvature:
        o.append(job.read(capi.OUT_TOTAL_CURVATURE))
    job.close()
    return o


CASES = [
    # dtype, integrator, curvature, kahan, tile bytes, halo, block, reorder, specialize
    (capi.F16, 1, 3, 0, 32768, 1, 256, True, 1),
    (capi.F32, 1, 3, 1, 32768, 1, 128, True, 1),
    (capi.BF16, 0, 0, 0, 49152, 2, 64, False, 1),
    (capi.F16, 1, 0, 0, 2048, 0, 256, True, 1),
    (capi.F32, 1, 3, 0, 61440, 1, 256, True, 0),
]


@pytest.mark.parametrize("dtype,integ,curv,kahan,tile,halo,block,reorder,spec", CASES)
def test_tile_equals_untiled(gpu_available, dtype, integ, curv, kahan, tile, halo, block, reorder, spec):
    src, gs = synthetic.abc_flow(24, 4)
    ctx = capi.Context(0)
    f = capi.Field(ctx, 24, 24, 24, 3, dtype, nslots=4)
    for sl in range(4):
        f.upload(sl, [src[sl, c] for c in range(3)])
    seeds = synthetic.uniform_seeds(8, 12000, [0.05] * 3, [2 * math.pi - 0.05] * 3)
    d = _desc(gs, 24, integ, curv, kahan)
    ref = _run(ctx, f, d, seeds, 2, 100, 0, block=block, reorder=reorder, specialize=spec)
    assert ctx.take_tile_stats() == (0, 0)
    got = _run(ctx, f, d, seeds, 2, 100, tile, halo=halo, block=block, reorder=reorder, specialize=spec)
    tiled, untiled = ctx.take_tile_stats()
    for a, b in zip(ref, got):
        np.testing.assert_array_equal(a, b)
    assert (ref[-4 if curv else -3] != 0).sum() > 50  # particles stopped along the way (3D border)
    if tile >= 16384 and reorder:
        assert tiled > 0  # cell-ordered workgroups staged tiles
    if tile == 2048:
        assert untiled > 0  # boxes that do not fit: the global fallback ran
    assert tiled + untiled > 0
    f.close()
    ctx.close()


@pytest.mark.parametrize("dtype", [capi.F16, capi.F32])
def test_tile_full_abc_field_equals_untiled(gpu_available, dtype):
    """The 3D ABC configuration's real 128^3 x 16 field, 65,536 of its seeds, RK4 dt 0.005 with
    curvature and sqrt(u*u+v*v+w*w) saved every step, through a window edge (200 + 100 steps)."""
    wind, gs = synthetic.abc_flow(128, 16)
    seeds = synthetic.uniform_seeds(1, 1 << 20, [0.0] * 3, [2 * math.pi] * 3)[::16].copy()
    ctx = capi.Context(0)
    f = capi.Field(ctx, 128, 128, 128, 3, dtype, nslots=3)
    for sl in range(3):
        f.upload(sl, [wind[sl, c] for c in range(3)])
    d = _desc(gs, 128, 1, 3, 0, expr="sqrt(u*u+v*v+w*w)", dt=0.005)
    outs = []
    for tile in (0, 32768):
        job = capi.Job(ctx, d, seeds, save_capacity=300)
        job.set_option(capi.OPT_QUAD, 0)
        job.set_option(capi.OPT_TILE, tile)
        job.reorder()
        t = np.float32(0)
        counter = 0
        for wb, steps in ((0, 200), (1, 100)):
            w = capi.Window()
            w.wind, w.wind_back, w.wind_front = f.h, wb, wb + 1
            w.time_back, w.time_front = float(wb), float(wb + 1)
            w.time_back64, w.time_front64 = float(wb), float(wb + 1)
            w.time_start = w.time_start64 = float(t)
            w.counter_start, w.nsteps = counter, steps
            job.advance(w)
            for _ in range(steps):
                t = np.float32(t + np.float32(0.005))
            counter += steps
        outs.append([job.read_saved(capi.OUT_POSITIONS, 0, 300), job.read_saved(capi.OUT_CURVATURE, 0, 300),
                     job.read_saved(capi.OUT_EXPR0, 0, 300), job.read(capi.OUT_STATUS),
                     job.read(capi.OUT_ABORT_STEP), job.read(capi.OUT_TOTAL_CURVATURE)])
        job.close()
    tiled, untiled = ctx.take_tile_stats()
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a, b)
    assert tiled > 0
    f.close()
    ctx.close()


def _windows(dt, total):
    """The reference's FP32 host schedule (a window edge once Time >= the window's end): the
    (window, steps, start time, first step) of each launch of a `total`-step run."""
    dt = np.float32(dt)
    t, mwt, counter, win, out = np.float32(0), np.float32(1), 0, 0, []
    while counter < total:
        t0, k, edge = t, 0, False
        while counter + k < total and not edge:
            k += 1
            t = np.float32(t + dt)
            edge = bool(t >= mwt)
        out.append((win, k, t0, counter))
        counter += k
        if edge:
            win += 1
            mwt = np.float32(mwt + np.float32(1))
    return out


def test_tile_matches_oracle(gpu_available):
    """The tile kernel against the FP32 oracle directly (positions, statuses, abort steps,
    curvature and the expression of every saved step), 3D ABC 32^3, 16,384 seeds, RK4 dt 0.01
    through a window edge."""
    wind, gs = synthetic.abc_flow(32, 4)
    seeds = synthetic.uniform_seeds(5, 16384, [0.0] * 3, [2 * math.pi] * 3)
    st = dict(Space=0, Integrator=1, DeltaT=0.01, DataTimeDistance=1.0, GridScale=gs, Offset=[0, 0, 0],
              MaximumSimulationTime=1.5, SaveInterval=1, CurvatureMode=3, AdditionalCompute=["sqrt(u*u+v*v+w*w)"])
    ref = oracle.run(st, wind, seeds, real64=False)
    total = int(ref["steps"])
    assert total >= 140
    ctx = capi.Context(0)
    f = capi.Field(ctx, 32, 32, 32, 3, capi.F32, nslots=4)
    for sl in range(4):
        f.upload(sl, [wind[sl, c] for c in range(3)])
    d = _desc(gs, 32, 1, 3, 0, expr="sqrt(u*u+v*v+w*w)")
    job = capi.Job(ctx, d, seeds, save_capacity=total)
    job.set_option(capi.OPT_QUAD, 0)
    job.set_option(capi.OPT_TILE, 61440)
    job.reorder()
    plan = _windows(0.01, total)
    assert len(plan) == 2
    for win, k, t0, first in plan:
        w = capi.Window()
        w.wind, w.wind_back, w.wind_front = f.h, win, win + 1
        w.time_back, w.time_front = float(np.float32(win)), float(np.float32(win + 1))
        w.time_back64, w.time_front64 = float(win), float(win + 1)
        w.time_start = w.time_start64 = float(t0)
        w.counter_start, w.nsteps = first, k
        job.advance(w)
    tiled, _ = ctx.take_tile_stats()
    assert tiled > 0
    gpu = {"trajectories": np.transpose(job.read_saved(capi.OUT_POSITIONS, 0, total), (1, 0, 2)),
           "curvatures": job.read_saved(capi.OUT_CURVATURE, 0, total).T[:, 1:],
           "status": job.read(capi.OUT_STATUS), "abort_step": job.read(capi.OUT_ABORT_STEP),
           "total_curvature": job.read(capi.OUT_TOTAL_CURVATURE)}
    expr = job.read_saved(capi.OUT_EXPR0, 0, total).T
    job.close()
    f.close()
    ctx.close()
    want = {"trajectories": ref["trajectories"][:, 1:], "curvatures": ref["curvatures"],
            "status": ref["status"], "abort_step": ref["abort_step"], "total_curvature": ref["total_curvature"]}
    assert_bitexact(gpu, want, fields=("trajectories", "curvatures", "total_curvature"), label="tile vs oracle ")
    np.testing.assert_array_equal(expr, ref["compute"][0].astype(np.float32))
