"""GPU parity for the kernel features outside the benchmark configurations: 2D loop borders,
sphere coordinates, Length / FitLength aborts, uplift, constants, additional volumes,
topography, reverse integration, constant wind fields, MinimumAliveParticles, SaveInterval 0,
and chunking invariance through the C ABI."""
import math

import numpy as np
import pytest

import oracle
from helpers import assert_bitexact, run_product
from pygranite_amd import capi, synthetic

pytestmark = pytest.mark.gpu


def gyre(n=64, m=32, S=4):
    wind, gs = synthetic.double_gyre(n, m, S)
    return wind, gs


def base2(gs, **kw):
    d = dict(Space="Space2D", Integrator="ClassicRungeKutta", DeltaT=0.02, DataTimeDistance=1.0,
             GridScale=gs, Offset=[0.0, 0.0], MaximumSimulationTime=2.0, SaveInterval=2)
    d.update(kw)
    return d


def abc(n=24, S=4):
    return synthetic.abc_flow(n, S)


def base3(gs, **kw):
    d = dict(Space="Space3D", Integrator="ClassicRungeKutta", DeltaT=0.01, DataTimeDistance=1.0,
             GridScale=gs, Offset=[0.0, 0.0, 0.0], MaximumSimulationTime=1.5, SaveInterval=3)
    d.update(kw)
    return d


def seeds2(n, seed=7):
    return synthetic.uniform_seeds(seed, n, [0.02, 0.02], [1.98, 0.98])


def seeds3(n, seed=8):
    return synthetic.uniform_seeds(seed, n, [0.05] * 3, [2 * math.pi - 0.05] * 3)


@pytest.mark.parametrize("border", ["LoopX", "LoopY", "LoopXY"])
def test_loop_borders_2d(pg, gpu_available, border):
    wind, gs = gyre()
    wind = wind.copy()
    wind[:, 0] += 0.3  # drift so particles cross the x edges
    wind[:, 1] += 0.2
    st = base2(gs, BorderMode=border, Integrator="ExplicitEuler")
    s = seeds2(1024)
    gpu = run_product(pg, st, wind, s)
    ref = oracle.run(st, wind, s, real64=False)
    assert_bitexact(gpu, ref, label=border + " ")


def test_sphere_coordinates_loopxy(pg, gpu_available):
    """Sphere coordinates call cosf (device libm vs glibc): compared at tolerance."""
    N, M = 72, 36
    lon = np.arange(N) * 5.0
    lat = np.arange(M) * 5.0 - 87.5
    wind = np.zeros((3, 2, M, N), np.float32)
    wind[:, 0] = 20.0
    wind[:, 1] = 5.0 * np.cos(np.radians(lon))[None, None, :]
    st = base2([5.0, 5.0], BorderMode="LoopXY", SphereRadius=6371000.0, DeltaT=600.0,
               DataTimeDistance=3600.0, MaximumSimulationTime=7200.0, Offset=[0.0, -90.0],
               SaveInterval=1)
    s = synthetic.uniform_seeds(9, 512, [0.0, -80.0], [355.0, 80.0])
    gpu = run_product(pg, st, wind, s)
    ref = oracle.run(st, wind, s, real64=False)
    assert np.array_equal(gpu["status"], ref["status"])
    np.testing.assert_allclose(gpu["trajectories"], ref["trajectories"], rtol=0, atol=2e-3)


@pytest.mark.parametrize("mode", ["Length", "FitLength"])
def test_length_aborts(pg, gpu_available, mode):
    wind, gs = abc()
    st = base3(gs, AbortMode=mode, MaximumLength=1.5)
    s = seeds3(1024)
    gpu = run_product(pg, st, wind, s)
    ref = oracle.run(st, wind, s, real64=False)
    assert_bitexact(gpu, ref, fields=("trajectories", "lengths"), label=mode + " ")
    assert (gpu["status"] == 1).sum() > 100


def test_reverse_and_constant_windfield(pg, gpu_available):
    wind, gs = abc()
    for extra in (dict(Reverse=True), dict(WindfieldMode="Constant")):
        st = base3(gs, **extra)
        s = seeds3(512)
        gpu = run_product(pg, st, wind, s)
        ref = oracle.run(st, wind, s, real64=False)
        assert_bitexact(gpu, ref, label=str(extra) + " ")


@pytest.mark.parametrize("modes", [dict(UpLiftMode="Dynamic", ConstantsMode="Dynamic", AdditionalVolumeMode="Constant"),
                                   dict(UpLiftMode="Constant", ConstantsMode="Constant", AdditionalVolumeMode="Dynamic")])
def test_uplift_constants_volumes_topography(pg, gpu_available, modes):
    wind, gs = abc(24, 4)
    n = 512
    s = seeds3(n)
    rng = np.random.default_rng(3)
    S = wind.shape[0]
    uplift = rng.normal(0, 1e-3, (S, n, 3)).astype(np.float32)
    consts = {"alpha": rng.uniform(0, 1, (S, n)).astype(np.float32),
              "beta": rng.uniform(1, 2, (S, n)).astype(np.float32)}
    vols = {"temp": rng.uniform(250, 300, (S, 24, 24, 24)).astype(np.float32),
            "qv": rng.uniform(0, 0.01, (S, 24, 24, 24)).astype(np.float32)}
    topo = np.full((24, 24), 0.3, np.float32)
    st = base3(gs, AdditionalCompute=["alpha*temp + beta*qv", "x+y+z-u*v"], **modes)
    gpu = run_product(pg, st, wind, s, uplift=uplift, constants=consts, volumes=vols,
                      topography=topo)
    ref = oracle.run(st, wind, s, real64=False, uplift=uplift, constants=consts, volumes=vols,
                     topography=topo)
    assert_bitexact(gpu, ref, label="extras ")
    for e in range(2):
        np.testing.assert_array_equal(gpu["compute"][e], ref["compute"][e].astype(np.float32))
    for k in vols:
        np.testing.assert_array_equal(gpu["volumes"][k], ref["volumes"][k].astype(np.float32))
    assert (gpu["status"] == 3).sum() > 0  # some particles hit the topography


def test_fast_total_curvature(pg, gpu_available):
    wind, gs = abc()
    st = base3(gs, CurvatureMode="FastTotalCurvature")
    s = seeds3(512)
    gpu = run_product(pg, st, wind, s)
    ref = oracle.run(st, wind, s, real64=False)
    assert_bitexact(gpu, ref, label="fast ")
    # acosf differs between device libm and glibc by ulps
    np.testing.assert_allclose(gpu["total_curvature"], ref["total_curvature"], rtol=1e-4, atol=1e-4)


def test_minimum_alive_and_save_interval_zero(pg, gpu_available):
    wind, gs = gyre()
    st = base2(gs, MinimumAliveParticles=400, SaveInterval=0, MaximumSimulationTime=3.0)
    s = synthetic.uniform_seeds(5, 512, [0.02, 0.02], [1.98, 0.98])
    wind = wind.copy()
    wind[:, 0] += 0.5  # particles leave through x = 2 over time
    gpu = run_product(pg, st, wind, s)
    ref = oracle.run(st, wind, s, real64=False)
    assert_bitexact(gpu, ref, label="alive ")
    assert gpu["trajectories"].shape[1] == 2


def test_chunked_advances_equal_one_advance(gpu_available):
    """Splitting a window into several launches gives identical state (C ABI)."""
    wind, gs = abc(32, 3)
    ctx = capi.Context(0)
    f = capi.Field(ctx, 32, 32, 32, 3, capi.F32, nslots=3)
    for sl in range(3):
        f.upload(sl, [wind[sl, c] for c in range(3)])
    s = seeds3(2048)
    outs = []
    for splits in ([100], [7, 13, 30, 50]):
        d = capi.JobDesc()
        d.dim, d.integrator, d.curvature, d.windfield_mode = 3, 1, 3, 1
        d.dt, d.dtd = 0.01, 1.0
        for k in range(3):
            d.grid_scale[k] = gs[k]
            d.domain[k] = 32 * gs[k]
        d.save_interval = 1
        d.nprog = 1
        d.prog[0] = capi.compile_expr("sqrt(u*u+v*v+w*w)", curvature=3)
        job = capi.Job(ctx, d, s, save_capacity=100)
        counter, t = 0, np.float32(0)
        for k in splits:
            w = capi.Window()
            w.wind, w.wind_back, w.wind_front = f.h, 0, 1
            w.time_back, w.time_front = 0.0, 1.0
            w.time_start, w.counter_start, w.nsteps = float(t), counter, k
            job.advance(w)
            for _ in range(k):
                t = np.float32(t + np.float32(0.01))
            counter += k
        outs.append((job.read(capi.OUT_POSITIONS), job.read_saved(capi.OUT_POSITIONS, 0, 100),
                     job.read_saved(capi.OUT_EXPR0, 0, 100), job.read(capi.OUT_TOTAL_CURVATURE)))
        job.close()
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a, b)
    f.close()
    ctx.close()


def test_saved_columns_of_stopped_particles(gpu_available):
    """The step kernels save only particles alive at the saving step; reading completes the
    columns with each stopped particle's final state. Checked against the FP32 oracle on a
    reordered job whose particles leave the domain at many different steps, with partial and
    repeated reads, and on the local-coordinate form (FP64 offset added on output)."""
    wind, gs = abc(24, 3)
    ctx = capi.Context(0)
    f = capi.Field(ctx, 24, 24, 24, 3, capi.F32, nslots=3)
    for sl in range(3):
        f.upload(sl, [wind[sl, c] for c in range(3)])
    s = synthetic.uniform_seeds(31, 5000, [0.0] * 3, [2 * math.pi] * 3)
    steps, si = 96, 3  # inside the first window (t < 1)
    st = base3(gs, DeltaT=0.01, MaximumSimulationTime=steps * 0.01, SaveInterval=si,
               CurvatureMode="IndividualAndTotalCurvature", AdditionalCompute=["sqrt(u*u+v*v+w*w)"])
    ref = oracle.run(st, wind, s, real64=False)
    assert (ref["status"] != 0).sum() > 1000  # many stop on the way, at different steps
    ncols = steps // si
    for local in (0, 1):
        d = capi.JobDesc()
        d.dim, d.integrator, d.curvature, d.windfield_mode = 3, 1, 3, 1
        d.dt, d.dtd = 0.01, 1.0
        for k in range(3):
            d.grid_scale[k] = gs[k]
            d.domain[k] = float(np.float32(24) * np.float32(gs[k]))
        d.save_interval = si
        d.local_coords = local
        d.nprog = 1
        d.prog[0] = capi.compile_expr("sqrt(u*u+v*v+w*w)", curvature=3)
        job = capi.Job(ctx, d, s, save_capacity=ncols)
        job.reorder()
        w = capi.Window()
        w.wind, w.wind_back, w.wind_front = f.h, 0, 1
        w.time_back, w.time_front = 0.0, 1.0
        w.time_start, w.counter_start, w.nsteps = 0.0, 0, steps
        job.advance(w)
        assert job.saved_columns() == ncols
        # partial reads, then the whole range again (the completion is idempotent)
        a = job.read_saved(capi.OUT_POSITIONS, 0, 20)
        b = job.read_saved(capi.OUT_POSITIONS, 20, ncols - 20)
        pos = job.read_saved(capi.OUT_POSITIONS, 0, ncols)
        np.testing.assert_array_equal(np.concatenate([a, b]), pos)
        want = ref["trajectories"][:, 1:ncols + 1, :].astype(np.float32)  # column k = step 3k
        got = np.transpose(pos.reshape(ncols, 5000, 3), (1, 0, 2))
        np.testing.assert_array_equal(got, want)
        ex = job.read_saved(capi.OUT_EXPR0, 0, ncols).reshape(ncols, 5000).T
        np.testing.assert_array_equal(ex, ref["compute"][0][:, :ncols].astype(np.float32))
        np.testing.assert_array_equal(job.read(capi.OUT_STATUS), ref["status"])
        job.close()
    f.close()
    ctx.close()


@pytest.mark.parametrize("dim,dtype,paired,curv,kahan", [
    (3, capi.F16, 1, 3, 0), (3, capi.F32, 0, 3, 0), (3, capi.BF16, 1, 0, 0), (2, capi.F32, 1, 0, 0),
    (2, capi.F16, 0, 3, 0), (3, capi.F32, 1, 3, 1), (3, capi.F16, 0, 0, 1), (2, capi.BF16, 1, 3, 1)])
def test_specialized_kernels_equal_generic(gpu_available, dim, dtype, paired, curv, kahan):
    """The step kernels built for the plain feature set (curvature Off / IndividualAndTotal,
    with or without Kahan-compensated positions) give every output bit for bit equal to the
    generic kernels (GR_OPT_SPECIALIZE = 0), over a launch that crosses window edges and saves
    every step."""
    ctx = capi.Context(0)
    if dim == 3:
        src, gs = abc(24, 4)
        f = capi.Field(ctx, 24, 24, 24, 3, dtype, nslots=4)
        seeds = seeds3(3000)
    else:
        src, gs = gyre(64, 32, 4)
        f = capi.Field(ctx, 64, 32, 1, 2, dtype, nslots=4, address=1)
        seeds = seeds2(3000)
    for sl in range(4):
        f.upload(sl, [src[sl, c] for c in range(dim)])
    outs = []
    for spec in (1, 0):
        d = capi.JobDesc()
        d.dim, d.integrator, d.curvature, d.windfield_mode = dim, 1, curv, 1
        d.dt, d.dtd = 0.01, 1.0
        d.kahan = kahan
        for k in range(3):
            d.grid_scale[k] = gs[k % len(gs)]
            d.domain[k] = float(np.float32(src.shape[-1 - k] if k < dim else 1) * np.float32(gs[k % len(gs)]))
        d.save_interval = 1
        d.nprog = 1
        d.prog[0] = capi.compile_expr("sqrt(u*u+v*v)*2-x", dim=dim, curvature=curv)
        job = capi.Job(ctx, d, seeds, save_capacity=250)
        job.set_option(capi.OPT_PAIRED, paired)
        job.set_option(capi.OPT_SPECIALIZE, spec)
        ends, _ = ctx.batch_run([(job, f, dict(max_steps=250))], ends=True)
        assert ends[0]["steps"] == 250
        o = [job.read_saved(capi.OUT_POSITIONS, 0, 250), job.read_saved(capi.OUT_EXPR0, 0, 250),
             job.read(capi.OUT_STATUS), job.read(capi.OUT_ABORT_STEP)]
        if curv:
            o += [job.read_saved(capi.OUT_CURVATURE, 0, 250), job.read(capi.OUT_TOTAL_CURVATURE)]
        # the single-job kernels from the same state: one more window through gr_job_advance
        w = capi.Window()
        w.wind, w.wind_back, w.wind_front = f.h, 2, 3
        w.time_back, w.time_front = 2.0, 3.0
        w.time_start = w.time_start64 = float(ends[0]["time"])
        w.counter_start, w.nsteps = 250, 40
        job.advance(w)
        o += [job.read(capi.OUT_POSITIONS), job.read(capi.OUT_EXPR0), job.read(capi.OUT_STATUS)]
        outs.append(o)
        job.close()
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a, b)
    if dim == 3:
        assert (outs[0][-4 if curv else -3] != 0).sum() > 50  # particles stopped along the way (3D border)
    f.close()
    ctx.close()


@pytest.mark.parametrize("reorder", [False, True])
def test_async_column_reads_overlapping_launches(gpu_available, reorder):
    """Saved columns read with gr_job_read_saved_async while the next chunk is already queued
    (a ring of two chunks: each launch overwrites the slots of the chunk before the one being
    read) equal the columns a twin job reads synchronously after each chunk."""
    wind, gs = abc(24, 4)
    ctx = capi.Context(0)
    f = capi.Field(ctx, 24, 24, 24, 3, capi.F32, nslots=4)
    for sl in range(4):
        f.upload(sl, [wind[sl, c] for c in range(3)])
    s = seeds3(5000, seed=12)
    chunk, nchunks = 20, 6

    def make():
        d = capi.JobDesc()
        d.dim, d.integrator, d.curvature, d.windfield_mode = 3, 1, 3, 1
        d.dt, d.dtd = 0.01, 1.0
        for k in range(3):
            d.grid_scale[k] = gs[k]
            d.domain[k] = float(np.float32(24) * np.float32(gs[k]))
        d.save_interval = 1
        d.nprog = 1
        d.prog[0] = capi.compile_expr("sqrt(u*u+v*v+w*w)", dim=3, curvature=3)
        j = capi.Job(ctx, d, s, save_capacity=2 * chunk)
        if reorder:
            j.reorder()
        return j

    a, b = make(), make()
    sync_cols, async_cols, pending = [], [], []
    counter, t = 0, np.float32(0)
    for k in range(nchunks):
        w = capi.Window()
        w.wind, w.wind_back, w.wind_front = f.h, 0, 1
        w.time_back, w.time_front = 0.0, 1.0
        w.time_start = w.time_start64 = float(t)
        w.counter_start, w.nsteps = counter, chunk
        a.advance(w)
        b.advance(w)
        for _ in range(chunk):
            t = np.float32(t + np.float32(0.01))
        first = counter
        counter += chunk
        sync_cols.append([b.read_saved(kind, first, chunk) for kind in (capi.OUT_POSITIONS, capi.OUT_CURVATURE, capi.OUT_EXPR0)])
        pending.append([a.read_saved_async(kind, first, chunk) for kind in (capi.OUT_POSITIONS, capi.OUT_CURVATURE, capi.OUT_EXPR0)])
        if len(pending) == 2:  # collect the previous chunk while this one computes
            done = pending.pop(0)
            for tk, _ in done:
                a.wait(tk)
            async_cols.append([arr for _, arr in done])
    for tk, _ in pending[0]:
        a.wait(tk)
    async_cols.append([arr for _, arr in pending[0]])
    assert len(async_cols) == nchunks
    for x, y in zip(sync_cols, async_cols):
        for u, v in zip(x, y):
            np.testing.assert_array_equal(u, v)
    a.close()
    b.close()
    f.close()
    ctx.close()


@pytest.mark.parametrize("dim,dtype,curv,kahan", [(3, capi.F32, 3, 0), (3, capi.F16, 3, 1), (3, capi.BF16, 0, 0),
                                                  (2, capi.F32, 0, 1), (2, capi.F16, 3, 0)])
def test_quad_launches_equal_one_lane(gpu_available, dim, dtype, curv, kahan):
    """Launches with four lanes per particle (GR_OPT_QUAD above the live range: every launch of
    the job) give every output bit for bit equal to one lane per particle (GR_OPT_QUAD = 0), over
    single-job advances through two windows of a regrouped job that save every step."""
    ctx = capi.Context(0)
    if dim == 3:
        src, gs = abc(24, 4)
        f = capi.Field(ctx, 24, 24, 24, 3, dtype, nslots=4)
        seeds = seeds3(3000)
    else:
        src, gs = gyre(64, 32, 4)
        f = capi.Field(ctx, 64, 32, 1, 2, dtype, nslots=4, address=1)
        seeds = seeds2(3000)
    for sl in range(4):
        f.upload(sl, [src[sl, c] for c in range(dim)])
    outs = []
    for quad in (0, 1 << 20):
        d = capi.JobDesc()
        d.dim, d.integrator, d.curvature, d.windfield_mode = dim, 1, curv, 1
        d.dt, d.dtd = 0.01, 1.0
        d.kahan = kahan
        for k in range(3):
            d.grid_scale[k] = gs[k % len(gs)]
            d.domain[k] = float(np.float32(src.shape[-1 - k] if k < dim else 1) * np.float32(gs[k % len(gs)]))
        d.save_interval = 1
        d.nprog = 1
        d.prog[0] = capi.compile_expr("sqrt(u*u+v*v)*2-x", dim=dim, curvature=curv)
        job = capi.Job(ctx, d, seeds, save_capacity=200)
        job.set_option(capi.OPT_QUAD, quad)
        job.reorder()
        counter, t = 0, np.float32(0)
        o = []
        for wb in (0, 1):
            w = capi.Window()
            w.wind, w.wind_back, w.wind_front = f.h, wb, wb + 1
            w.time_back, w.time_front = float(wb), float(wb + 1)
            w.time_start = w.time_start64 = float(t)
            w.counter_start, w.nsteps = counter, 100
            job.advance(w)
            # this window's columns, read before the particles are regrouped again
            o += [job.read_saved(capi.OUT_POSITIONS, counter, 100), job.read_saved(capi.OUT_EXPR0, counter, 100)]
            if curv:
                o.append(job.read_saved(capi.OUT_CURVATURE, counter, 100))
            for _ in range(100):
                t = np.float32(t + np.float32(0.01))
            counter += 100
            job.reorder()
        o += [job.read(capi.OUT_STATUS), job.read(capi.OUT_ABORT_STEP), job.read(capi.OUT_POSITIONS)]
        if curv:
            o.append(job.read(capi.OUT_TOTAL_CURVATURE))
        outs.append(o)
        job.close()
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a, b)
    if dim == 3:
        assert (outs[0][-4 if curv else -3] != 0).sum() > 50  # particles stopped along the way (3D border)
    f.close()
    ctx.close()
