"""The persistent batched kernel: jobs of mixed dimension, integrator and storage type in one
launch, each crossing several windows, compared bit for bit with the FP32 oracle (fed the same
rounded slices) and with the same job run alone."""
import math

import numpy as np
import pytest

import oracle
from pygranite_amd import capi, synthetic

pytestmark = pytest.mark.gpu

ROUND = {capi.F32: lambda a: a, capi.F16: synthetic.round_fp16, capi.BF16: synthetic.round_bf16}


def make_desc(dim, integ, gs, shape, dt, curvature=0, exprs=(), save_interval=0):
    d = capi.JobDesc()
    d.dim, d.integrator, d.curvature, d.windfield_mode = dim, integ, curvature, 1
    d.dt, d.dtd = dt, 1.0
    n = [shape[-1], shape[-2], shape[-3] if dim == 3 else 1]
    for k in range(3):
        d.grid_scale[k] = gs[k] if k < len(gs) else gs[0]
        d.domain[k] = float(np.float32(n[k]) * np.float32(d.grid_scale[k]))
    d.save_interval = save_interval
    d.nprog = len(exprs)
    for i, e in enumerate(exprs):
        d.prog[i] = capi.compile_expr(e, dim=dim, curvature=curvature)
    return d


def oracle_settings(dim, integ, gs, dt, T, curvature=0, exprs=()):
    return dict(Space=0 if dim == 3 else 1, Integrator=integ, DeltaT=dt, DataTimeDistance=1.0,
                GridScale=list(gs), Offset=[0.0] * dim, MaximumSimulationTime=T, SaveInterval=0,
                CurvatureMode=curvature, AdditionalCompute=list(exprs))


JOBS = [
    # dim, integrator, dtype, particles, simulated time, curvature, expressions
    (2, 0, capi.F16, 700, 2.5, 0, ()),
    (2, 1, capi.F32, 1500, 3.2, 0, ("sqrt(u*u+v*v)",)),
    (3, 1, capi.F16, 2048, 2.2, 3, ("sqrt(u*u+v*v+w*w)",)),
    (3, 0, capi.BF16, 333, 1.7, 0, ()),
    (3, 1, capi.F32, 1000, 3.6, 2, ()),
    (2, 1, capi.BF16, 64, 0.5, 2, ()),
]


@pytest.mark.parametrize("quad", [True, False])
def test_batch_mixed_jobs_bitexact(gpu_available, quad):
    """Every job of the batch equals the FP32 oracle, with quad items (the launch's few live
    particles: a quad of lanes per particle for the specialised jobs) and without (quad limit 0)."""
    ctx = capi.Context(0)
    g2, gs2 = synthetic.double_gyre(96, 48, 6)
    g3, gs3 = synthetic.abc_flow(24, 6)
    fields = {}
    for dim, src in ((2, g2), (3, g3)):
        for dt in (capi.F32, capi.F16, capi.BF16):
            S = src.shape[0]
            shp = src.shape[2:]
            f = capi.Field(ctx, shp[-1], shp[-2], shp[-3] if dim == 3 else 1, dim, dt, nslots=S,
                           address=0 if dim == 3 else 1)
            for s in range(S):
                f.upload(s, [src[s, c] for c in range(dim)])
            fields[(dim, dt)] = f
    jobs, spans, refs = [], [], []
    for idx, (dim, integ, dt_, n, T, curv, exprs) in enumerate(JOBS):
        src, gs = (g2, gs2) if dim == 2 else (g3, gs3)
        step = 0.01 if dim == 2 else 0.005
        lo = [0.02] * dim
        hi = [1.98, 0.98] if dim == 2 else [2 * math.pi - 0.02] * 3
        seeds = synthetic.uniform_seeds(100 + idx, n, lo, hi)
        d = make_desc(dim, integ, gs, src.shape, step, curv, exprs)
        job = capi.Job(ctx, d, seeds)
        if not quad:
            job.set_option(capi.OPT_QUAD, 0)
        if idx % 2 == 0:
            job.reorder()  # batched items read reordered state too
        jobs.append(job)
        spans.append((job, fields[(dim, dt_)], dict(max_steps=1 << 30, max_sim_time=T)))
        st = oracle_settings(dim, integ, gs, step, T, curv, exprs)
        refs.append(oracle.run(st, ROUND[dt_](src), seeds, real64=False))
    steps, ms = ctx.batch_run(spans)
    assert ms > 0
    # quad items for the four specialised jobs (curvature 0 or 3): ceil(n / 16) each
    want_quad = sum(-(-n // 16) for (_, _, _, n, _, curv, _) in JOBS if curv in (0, 3)) if quad else 0
    assert ctx.get_option(capi.CTX_BATCH_QUAD_ITEMS) == want_quad
    for idx, (job, ref) in enumerate(zip(jobs, refs)):
        assert steps[idx] == ref["steps"], (idx, steps[idx], ref["steps"])
        pos = job.read(capi.OUT_POSITIONS)
        want = ref["trajectories"][:, -1, :].astype(np.float32)
        np.testing.assert_array_equal(pos, want, err_msg=f"job {idx}")
        np.testing.assert_array_equal(job.read(capi.OUT_STATUS), ref["status"])
        np.testing.assert_array_equal(job.read(capi.OUT_ABORT_STEP), ref["abort_step"])
        if JOBS[idx][6]:
            np.testing.assert_array_equal(job.read(capi.OUT_EXPR0),
                                          ref["compute"][0][:, -1].astype(np.float32))
    for j in jobs:
        j.close()
    for f in fields.values():
        f.close()
    ctx.close()


def test_batch_equals_single_job_advances(gpu_available):
    """A span in the batch (quad items) == the same job advanced window by window with
    gr_job_advance at one lane per particle."""
    ctx = capi.Context(0)
    src, gs = synthetic.abc_flow(32, 5)
    f = capi.Field(ctx, 32, 32, 32, 3, capi.F16, nslots=5)
    for s in range(5):
        f.upload(s, [src[s, c] for c in range(3)])
    seeds = synthetic.uniform_seeds(7, 4096, [0.1] * 3, [6.1] * 3)
    d = make_desc(3, 1, gs, src.shape, 0.005, 3, ("sqrt(u*u+v*v+w*w)",), save_interval=10)
    a = capi.Job(ctx, d, seeds, save_capacity=200)
    b = capi.Job(ctx, d, seeds, save_capacity=200)
    b.set_option(capi.OPT_QUAD, 0)  # one lane per particle
    steps, _ = ctx.batch_run([(a, f, dict(max_steps=700))])
    assert steps == [700]
    assert ctx.get_option(capi.CTX_BATCH_QUAD_ITEMS) == 4096 // 16  # a quad of lanes per particle
    # b: window by window under the same FP32 schedule
    t, counter, c, mwt, left = np.float32(0), 0, 0, np.float32(1.0), 700
    while left:
        k, tt, edge = 0, t, False
        while k < left:
            k += 1
            tt = np.float32(tt + np.float32(0.005))
            if tt >= mwt:
                edge = True
                break
        w = capi.Window()
        w.wind, w.wind_back, w.wind_front = f.h, c, c + 1
        w.time_back, w.time_front = float(np.float32(c)), float(np.float32(c + 1))
        w.time_start, w.counter_start, w.nsteps = float(t), counter, k
        b.advance(w)
        t, counter, left = tt, counter + k, left - k
        if edge:
            c += 1
            mwt = np.float32(mwt + np.float32(1.0))
    for kind in (capi.OUT_POSITIONS, capi.OUT_STATUS, capi.OUT_TOTAL_CURVATURE, capi.OUT_EXPR0):
        np.testing.assert_array_equal(a.read(kind), b.read(kind))
    assert a.saved_columns() == b.saved_columns() == 70
    np.testing.assert_array_equal(a.read_saved(capi.OUT_POSITIONS, 0, 70),
                                  b.read_saved(capi.OUT_POSITIONS, 0, 70))
    a.close()
    b.close()
    f.close()
    ctx.close()


@pytest.mark.parametrize("block,waves,group,cuw,spec", [(64, 0, 1, 6, 1), (128, 4, 8, 4, 1), (256, 4, 0, 0, 0),
                                                  (192, 0, 3, 8, 0), (64, 0, 0, 0, 0)])
def test_batch_launch_options_change_nothing(gpu_available, block, waves, group, cuw, spec):
    """Workgroup size, register budget, wavefronts per CU and the compile-time feature sets
    (spec 0: the generic kernels) of the batched kernel leave every result unchanged, and the
    device work counter equals the integrated particle-steps implied by the final status and
    abort step of every particle."""
    ctx = capi.Context(0)
    src, gs = synthetic.abc_flow(32, 5)
    f = capi.Field(ctx, 32, 32, 32, 3, capi.F16, nslots=5)
    for s in range(5):
        f.upload(s, [src[s, c] for c in range(3)])
    g2, gs2 = synthetic.double_gyre(96, 48, 6)
    f2 = capi.Field(ctx, 96, 48, 1, 2, capi.F32, nslots=6, address=1)
    for s in range(6):
        f2.upload(s, [g2[s, c] for c in range(2)])
    # seeds over the whole domain: most leave it (3D border) at different steps
    s3 = synthetic.uniform_seeds(17, 5000, [0.0] * 3, [2 * math.pi] * 3)
    s2 = synthetic.uniform_seeds(18, 3001, [0.02, 0.02], [1.98, 0.98])
    d3 = make_desc(3, 1, gs, src.shape, 0.005, 3, ("sqrt(u*u+v*v+w*w)",))
    d2 = make_desc(2, 0, gs2, g2.shape, 0.01)

    def run(opts):
        if opts:
            ctx.set_option(capi.CTX_BATCH_BLOCK, opts[0])
            ctx.set_option(capi.CTX_BATCH_WAVES, opts[1])
            if opts[2]:
                ctx.set_option(capi.CTX_BATCH_GROUP, opts[2])
            ctx.set_option(capi.CTX_BATCH_CU_WAVES, opts[3])
        a = capi.Job(ctx, d3, s3)
        a.reorder()
        b = capi.Job(ctx, d2, s2)
        if opts:
            a.set_option(capi.OPT_SPECIALIZE, opts[4])
            b.set_option(capi.OPT_SPECIALIZE, opts[4])
        ctx.take_work()
        steps, _ = ctx.batch_run([(a, f, dict(max_steps=640)), (b, f2, dict(max_steps=333))])
        work = ctx.take_work()
        out = [(j.read(capi.OUT_POSITIONS), j.read(capi.OUT_STATUS), j.read(capi.OUT_ABORT_STEP))
               for j in (a, b)]
        if len(out[0]) == 3:
            out[0] = out[0] + (a.read(capi.OUT_EXPR0), a.read(capi.OUT_TOTAL_CURVATURE))
        a.close()
        b.close()
        return steps, work, out

    steps0, work0, ref = run(None)
    assert ctx.get_option(capi.CTX_BATCH_BLOCK) == 256 and ctx.get_option(capi.CTX_BATCH_WAVES) == 0
    steps1, work1, got = run((block, waves, group, cuw, spec))
    assert steps0 == steps1 == [640, 333]
    for r, g in zip(ref, got):
        for x, y in zip(r, g):
            np.testing.assert_array_equal(x, y)
    expect = 0
    for (pos, status, ab), k in zip([r[:3] for r in ref], steps0):
        expect += int(np.where(status == 0, k, ab.astype(np.int64) + 1).sum())
    assert work0 == work1 == expect
    assert (ref[0][1] != 0).sum() > 100  # the 3D job really lost particles on the way
    with pytest.raises(capi.GraniteError):
        ctx.set_option(capi.CTX_BATCH_BLOCK, 96)
    with pytest.raises(capi.GraniteError):
        ctx.set_option(capi.CTX_BATCH_WAVES, 3)
    with pytest.raises(capi.GraniteError):
        ctx.set_option(capi.CTX_BATCH_GROUP, 0)
    f.close()
    f2.close()
    ctx.close()


def test_single_job_work_counter(gpu_available):
    """gr_job_advance adds the same integrated particle-steps to the context's work counter."""
    ctx = capi.Context(0)
    src, gs = synthetic.abc_flow(24, 3)
    f = capi.Field(ctx, 24, 24, 24, 3, capi.F32, nslots=3)
    for s in range(3):
        f.upload(s, [src[s, c] for c in range(3)])
    seeds = synthetic.uniform_seeds(19, 777, [0.0] * 3, [2 * math.pi] * 3)
    j = capi.Job(ctx, make_desc(3, 1, gs, src.shape, 0.005), seeds)
    w = capi.Window()
    w.wind, w.wind_back, w.wind_front = f.h, 0, 1
    w.time_back, w.time_front, w.time_start, w.counter_start, w.nsteps = 0.0, 1.0, 0.0, 0, 150
    ctx.take_work()
    j.advance(w)
    status, ab = j.read(capi.OUT_STATUS), j.read(capi.OUT_ABORT_STEP)
    assert ctx.take_work() == int(np.where(status == 0, 150, ab.astype(np.int64) + 1).sum())
    assert ctx.take_work() == 0
    j.close()
    f.close()
    ctx.close()


def test_autotune_picks_a_candidate_and_leaves_counters(gpu_available):
    """The autotuner times every candidate, sets the fastest on the context, and leaves the
    context's work counter and kernel-time sums as they were; later batches give the same
    results as with the default configuration."""
    ctx = capi.Context(0)
    src, gs = synthetic.abc_flow(32, 5)
    f = capi.Field(ctx, 32, 32, 32, 3, capi.F16, nslots=5)
    for s in range(5):
        f.upload(s, [src[s, c] for c in range(3)])
    d = make_desc(3, 1, gs, src.shape, 0.005, 3, ("sqrt(u*u+v*v+w*w)",))
    seeds = synthetic.uniform_seeds(23, 3000, [0.0] * 3, [2 * math.pi] * 3)
    a = capi.Job(ctx, d, seeds)
    ctx.batch_run([(a, f, dict(max_steps=100))])
    ctx.synchronize()
    res = ctx.autotune(f, d, n=8192, njobs=3, steps=40, trials=1)
    assert len(res) >= 6 and all(r["ms"] > 0 for r in res)
    # workgroup size x wavefronts per CU (register-limited, 4, 6, 8; a cap admits one workgroup)
    want = {(b_, c_) for b_ in (64, 128, 256) for c_ in (0, 4, 6, 8) if c_ == 0 or c_ >= b_ // 64}
    assert {(r["block"], r["cu_waves"]) for r in res} == want
    best = min(res, key=lambda r: r["ms"])
    assert ctx.get_option(capi.CTX_BATCH_BLOCK) == best["block"]
    assert ctx.get_option(capi.CTX_BATCH_WAVES) == best["waves"]
    assert ctx.get_option(capi.CTX_BATCH_GROUP) == best["group"]
    assert ctx.get_option(capi.CTX_BATCH_CU_WAVES) == best["cu_waves"]
    # the probe runs did not count: only job a's 100 steps are in the counters
    status, ab = a.read(capi.OUT_STATUS), a.read(capi.OUT_ABORT_STEP)
    assert ctx.take_work() == int(np.where(status == 0, 100, ab.astype(np.int64) + 1).sum())
    ms, launches = ctx.take_kernel_ms()
    assert launches == 1 and ms > 0
    # same results under the tuned configuration as under the default one
    b = capi.Job(ctx, d, seeds)
    ctx.batch_run([(b, f, dict(max_steps=100))])
    np.testing.assert_array_equal(a.read(capi.OUT_POSITIONS), b.read(capi.OUT_POSITIONS))
    a.close()
    b.close()
    f.close()
    ctx.close()


def test_joint_reorder_of_mixed_jobs(gpu_available):
    """Reordering 2D and 3D jobs together (one key kernel, one sort, one gather) changes no
    result: positions, status, abort steps and saved columns equal those of unreordered jobs,
    also after a second joint reorder half way."""
    ctx = capi.Context(0)
    src, gs = synthetic.abc_flow(32, 5)
    f = capi.Field(ctx, 32, 32, 32, 3, capi.F16, nslots=5)
    for s in range(5):
        f.upload(s, [src[s, c] for c in range(3)])
    g2, gs2 = synthetic.double_gyre(96, 48, 6)
    f2 = capi.Field(ctx, 96, 48, 1, 2, capi.F32, nslots=6, address=1)
    for s in range(6):
        f2.upload(s, [g2[s, c] for c in range(2)])
    d3 = make_desc(3, 1, gs, src.shape, 0.005, 3, ("sqrt(u*u+v*v+w*w)",), save_interval=7)
    d2 = make_desc(2, 1, gs2, g2.shape, 0.01, 0, (), save_interval=7)
    seeds = [synthetic.uniform_seeds(40, 9000, [0.0] * 3, [2 * math.pi] * 3),
             synthetic.uniform_seeds(41, 5000, [0.02, 0.02], [1.98, 0.98]),
             synthetic.uniform_seeds(42, 777, [0.0] * 3, [2 * math.pi] * 3)]
    descs, fields = [d3, d2, d3], [f, f2, f]

    def run(joint):
        jobs = [capi.Job(ctx, d, sd, save_capacity=40) for d, sd in zip(descs, seeds)]
        if joint:
            ctx.reorder(jobs)
        ends, _ = ctx.batch_run([(j, fl, dict(max_steps=140)) for j, fl in zip(jobs, fields)], ends=True)
        cols = [j.read_saved(capi.OUT_POSITIONS, 0, 20) for j in jobs]
        if joint:
            ctx.reorder(jobs)
        ctx.batch_run([(j, fl, dict(max_steps=140, first_window=e["window"], time_start=e["time"],
                                    time_start64=e["time64"], counter_start=e["counter"]))
                       for j, fl, e in zip(jobs, fields, ends)])
        out = [(j.read(capi.OUT_POSITIONS), j.read(capi.OUT_STATUS), j.read(capi.OUT_ABORT_STEP),
                j.read_saved(capi.OUT_POSITIONS, 20, 20)) for j in jobs]
        for j in jobs:
            j.close()
        return cols, out

    c0, o0 = run(False)
    c1, o1 = run(True)
    for a, b in zip(c0, c1):
        np.testing.assert_array_equal(a, b)
    for x, y in zip(o0, o1):
        for a, b in zip(x, y):
            np.testing.assert_array_equal(a, b)
    f.close()
    f2.close()
    ctx.close()



def _advance_chunks(job, f, steps, chunk, dt, after_chunk):
    """Advances `job` by `steps` in launches of at most `chunk` steps under the FP32 window
    schedule (a launch also ends at a window edge); after_chunk(steps_left) runs after each."""
    t, counter, c, mwt, left = np.float32(0), 0, 0, np.float32(1.0), steps
    launches = 0
    while left:
        k, tt, edge = 0, t, False
        while k < min(left, chunk):
            k += 1
            tt = np.float32(tt + np.float32(dt))
            if tt >= mwt:
                edge = True
                break
        w = capi.Window()
        w.wind, w.wind_back, w.wind_front = f.h, c, c + 1
        w.time_back, w.time_front = float(np.float32(c)), float(np.float32(c + 1))
        w.time_start, w.counter_start, w.nsteps = float(t), counter, k
        job.advance(w)
        t, counter, left = tt, counter + k, left - k
        if edge:
            c += 1
            mwt = np.float32(mwt + np.float32(1.0))
        launches += 1
        after_chunk(left)
    return launches


def test_live_range_bounds_launches(gpu_available):
    """Launches that cover only a job's live range (the leading slots a reorder filled with its
    alive particles) give the results of launches over every particle, on the single-job and
    the batched path; the range shrinks once a reorder's count has been taken up and always
    holds every alive particle."""
    ctx = capi.Context(0)
    src, gs = synthetic.abc_flow(32, 5)
    f = capi.Field(ctx, 32, 32, 32, 3, capi.F16, nslots=5)
    for s in range(5):
        f.upload(s, [src[s, c] for c in range(3)])
    # dt = 2^-7: FP32 time is exact, so window edges fall at steps 128, 256, 384 and every launch
    # has 32 steps; the last launch saves the columns of counters 450 ... 475
    n, dt, steps, chunk = 6000, 0.0078125, 480, 32
    seeds = [synthetic.uniform_seeds(9, n, [0.0] * 3, [2 * math.pi] * 3),
             synthetic.uniform_seeds(10, 2500, [0.0] * 3, [2 * math.pi] * 3)]
    d = make_desc(3, 1, gs, src.shape, dt, 3, ("sqrt(u*u+v*v+w*w)",), save_interval=5)
    kinds = (capi.OUT_POSITIONS, capi.OUT_STATUS, capi.OUT_ABORT_STEP, capi.OUT_TOTAL_CURVATURE,
             capi.OUT_EXPR0)

    def single(reorder):
        job = capi.Job(ctx, d, seeds[0], save_capacity=100)
        ranges = []

        def after(left):
            if reorder and left:
                job.reorder()
                ranges.append((job.live_range(), job.count_alive()))

        assert _advance_chunks(job, f, steps, chunk, dt, after) == steps // chunk
        out = [job.read(k) for k in kinds]
        out.append(job.read_saved(capi.OUT_POSITIONS, job.saved_columns() - 6, 6))
        job.close()
        return out, ranges

    def batched(reorder):
        jobs = [capi.Job(ctx, d, sd, save_capacity=100) for sd in seeds]
        ranges = []
        e = dict(window=0, time=0.0, time64=0.0, counter=0)
        for i in range(steps // chunk):
            ends, _ = ctx.batch_run([(j, f, dict(max_steps=chunk, first_window=e["window"], time_start=e["time"],
                                                 time_start64=e["time64"], counter_start=e["counter"]))
                                     for j in jobs], ends=True)
            e = ends[0]
            if reorder and i < steps // chunk - 1:
                ctx.reorder(jobs)
                ranges.append([(j.live_range(), j.count_alive()) for j in jobs])
        assert e["counter"] == steps
        out = [[j.read(k) for k in kinds] + [j.read_saved(capi.OUT_POSITIONS, j.saved_columns() - 6, 6)]
               for j in jobs]
        for j in jobs:
            j.close()
        return out, ranges

    a, _ = single(False)
    b, ranges = single(True)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert all(c <= r <= n for r, c in ranges)
    # from the second reorder on, the first one's count has been taken up
    assert ranges[-1][0] <= ranges[1][0] < n

    a, _ = batched(False)
    b, ranges = batched(True)
    for ja, jb in zip(a, b):
        for x, y in zip(ja, jb):
            np.testing.assert_array_equal(x, y)
    for step in ranges:
        for (r, c), sd in zip(step, seeds):
            assert c <= r <= sd.shape[0]
    assert all(r < sd.shape[0] for (r, _), sd in zip(ranges[-1], seeds))
    f.close()
    ctx.close()


def test_wide_launch_register_budget_bitexact(gpu_available):
    """A launch of many more wavefronts than the device holds at once (here 400,000 particles)
    runs the FP16 time-paired kernel built for four wavefronts per SIMD; its results equal the
    batched kernel's bit for bit, and a subsample equals the FP32 oracle fed the same rounded
    slices."""
    ctx = capi.Context(0)
    src, gs = synthetic.abc_flow(32, 3)
    f = capi.Field(ctx, 32, 32, 32, 3, capi.F16, nslots=3)
    for s in range(3):
        f.upload(s, [src[s, c] for c in range(3)])
    n, dt = 400000, 0.0078125
    seeds = synthetic.uniform_seeds(12, n, [0.0] * 3, [2 * math.pi] * 3)
    d = make_desc(3, 1, gs, src.shape, dt, 3, ("sqrt(u*u+v*v+w*w)",))
    a = capi.Job(ctx, d, seeds)
    b = capi.Job(ctx, d, seeds)
    assert _advance_chunks(a, f, 48, 16, dt, lambda left: None) == 3
    steps, _ = ctx.batch_run([(b, f, dict(max_steps=48))])
    assert steps == [48]
    kinds = (capi.OUT_POSITIONS, capi.OUT_STATUS, capi.OUT_ABORT_STEP, capi.OUT_TOTAL_CURVATURE,
             capi.OUT_EXPR0)
    got = {k: a.read(k) for k in kinds}
    for k in kinds:
        np.testing.assert_array_equal(got[k], b.read(k))
    sub = np.arange(0, n, 200)
    st = oracle_settings(3, 1, gs, dt, 47.5 * dt, 3, ("sqrt(u*u+v*v+w*w)",))
    ref = oracle.run(st, synthetic.round_fp16(src), seeds[sub], real64=False)
    assert ref["steps"] == 48
    np.testing.assert_array_equal(got[capi.OUT_POSITIONS][sub], ref["trajectories"][:, -1].astype(np.float32))
    np.testing.assert_array_equal(got[capi.OUT_STATUS][sub], ref["status"])
    a.close()
    b.close()
    f.close()
    ctx.close()


def test_reorder_with_wide_keys_changes_nothing(gpu_available):
    """Three jobs on a field 1024 cells long: 30-bit Morton codes, of which the reorder's
    counting sort keeps the top 18 bits per job (a coarser regrouping than on small grids). The
    results equal those of unreordered jobs, and the live ranges hold every alive particle."""
    ctx = capi.Context(0)
    nx, ny, nz, S = 1024, 4, 4, 3
    x = np.arange(nx, dtype=np.float32) * np.float32(0.01)
    wind = np.zeros((S, 3, nz, ny, nx), np.float32)
    for s in range(S):
        wind[s, 0] = (0.6 + 0.3 * np.sin(7.0 * x + s))[None, None, :]
        wind[s, 1] = 0.05 * np.cos(3.0 * x)[None, None, :]
        wind[s, 2] = 0.02
    gs = [0.01, 0.01, 0.01]
    f = capi.Field(ctx, nx, ny, nz, 3, capi.F16, nslots=S)
    for s in range(S):
        f.upload(s, [wind[s, c] for c in range(3)])
    d = make_desc(3, 1, gs, wind.shape, 0.0078125, 0, (), save_interval=0)
    seeds = [synthetic.uniform_seeds(50 + k, 3000, [0.0, 0.0, 0.0], [10.2, 0.03, 0.03]) for k in range(3)]

    def run(joint):
        jobs = [capi.Job(ctx, d, sd) for sd in seeds]
        if joint:
            ctx.reorder(jobs)
        e = dict(window=0, time=0.0, time64=0.0, counter=0)
        ranges = []
        for _ in range(4):
            ends, _ = ctx.batch_run([(j, f, dict(max_steps=32, first_window=e["window"], time_start=e["time"],
                                                 time_start64=e["time64"], counter_start=e["counter"]))
                                     for j in jobs], ends=True)
            e = ends[0]
            if joint:
                ctx.reorder(jobs)
                ranges.append([(j.live_range(), j.count_alive()) for j in jobs])
        out = [(j.read(capi.OUT_POSITIONS), j.read(capi.OUT_STATUS), j.read(capi.OUT_ABORT_STEP)) for j in jobs]
        for j in jobs:
            j.close()
        return out, ranges

    a, _ = run(False)
    b, ranges = run(True)
    for x_, y_ in zip(a, b):
        for p, q in zip(x_, y_):
            np.testing.assert_array_equal(p, q)
    for step in ranges:
        for (r, c), sd in zip(step, seeds):
            assert c <= r <= sd.shape[0]
    assert any((s != 0).any() for _, s, _ in b)  # particles stopped (z or y edges), ranges shrank
    f.close()
    ctx.close()


def test_new_job_in_a_closed_jobs_place_gets_its_own_items(gpu_available):
    """A batched launch reuses its item list when its spans are unchanged. A job created after
    another was closed (and possibly at the same address) with the same particle count and field
    but a different integrator or dimension must not run the closed job's item variant: an RK4
    job after an Euler job of the same size equals the RK4 job run alone on a fresh context."""
    wind, gs = synthetic.abc_flow(24, 4)
    seeds = synthetic.uniform_seeds(31, 3000, [0.05] * 3, [2 * math.pi - 0.05] * 3)

    def field(ctx):
        f = capi.Field(ctx, 24, 24, 24, 3, capi.F32, nslots=4)
        for sl in range(4):
            f.upload(sl, [wind[sl, c] for c in range(3)])
        return f

    ctx = capi.Context(0)
    f = field(ctx)
    for integ in (0, 1, 0, 1):
        job = capi.Job(ctx, make_desc(3, integ, gs, (24, 24, 24), 0.01), seeds)
        ctx.batch_run([(job, f, dict(max_steps=60))], timed=False)
        pos = job.read(capi.OUT_POSITIONS)
        job.close()
    f.close()
    ctx.close()
    ctx2 = capi.Context(0)
    f2 = field(ctx2)
    ref = capi.Job(ctx2, make_desc(3, 1, gs, (24, 24, 24), 0.01), seeds)
    ctx2.batch_run([(ref, f2, dict(max_steps=60))], timed=False)
    want = ref.read(capi.OUT_POSITIONS)
    ref.close()
    f2.close()
    ctx2.close()
    np.testing.assert_array_equal(pos, want)


def test_single_job_autotune_sets_a_candidate_and_changes_nothing(gpu_available):
    """gr_job_autotune times every single-job launch configuration (workgroup size x lanes per
    particle x LDS field tile) on a probe job, sets the fastest on the job and leaves the job's
    state, the context's work counter and its tile statistics as they were; a tuned job's
    results equal an untuned twin's bit for bit. Jobs with per-particle data are refused."""
    wind, gs = synthetic.abc_flow(24, 4)
    ctx = capi.Context(0)
    f = capi.Field(ctx, 24, 24, 24, 3, capi.F16, nslots=4)
    for sl in range(4):
        f.upload(sl, [wind[sl, c] for c in range(3)])
    seeds = synthetic.uniform_seeds(41, 12000, [0.05] * 3, [2 * math.pi - 0.05] * 3)
    d = make_desc(3, 1, gs, (24, 24, 24), 0.01, curvature=3, exprs=("sqrt(u*u+v*v+w*w)",), save_interval=1)
    tuned, twin = (capi.Job(ctx, d, seeds, save_capacity=100) for _ in range(2))
    ctx.take_work()
    ctx.take_tile_stats()
    cands = tuned.autotune(f, steps=40, trials=1)
    assert len(cands) == 12  # 3 workgroup sizes x (4 lanes, or 1 lane with tiles of 0 / 16 / 32 KiB)
    assert {(c["block"], c["lanes"], c["tile_bytes"]) for c in cands} == \
        {(b, 4, 0) for b in (64, 128, 256)} | {(b, 1, t) for b in (64, 128, 256) for t in (0, 16384, 32768)}
    assert all(c["ms"] > 0 for c in cands)
    assert ctx.take_work() == 0 and ctx.take_tile_stats() == (0, 0)
    outs = []
    for j in (tuned, twin):
        j.reorder()
        w = capi.Window()
        w.wind, w.wind_back, w.wind_front = f.h, 0, 1
        w.time_back, w.time_front, w.time_back64, w.time_front64 = 0.0, 1.0, 0.0, 1.0
        w.time_start = w.time_start64 = 0.0
        w.counter_start, w.nsteps = 0, 100
        j.advance(w)
        outs.append([j.read_saved(capi.OUT_POSITIONS, 0, 100), j.read_saved(capi.OUT_EXPR0, 0, 100),
                     j.read(capi.OUT_STATUS), j.read(capi.OUT_TOTAL_CURVATURE)])
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a, b)
    dv = make_desc(3, 1, gs, (24, 24, 24), 0.01)
    dv.uplift_mode = 1
    withlift = capi.Job(ctx, dv, seeds[:100])
    with pytest.raises(capi.GraniteError):
        withlift.autotune(f, steps=10, trials=1)
    for j in (tuned, twin, withlift):
        j.close()
    f.close()
    ctx.close()


@pytest.mark.parametrize("dtype,kahan", [(capi.F32, 1), (capi.F16, 0), (capi.BF16, 1)])
def test_batch_quad_items_equal_one_lane_items(gpu_available, dtype, kahan):
    """Quad items (a quad of lanes per particle) against one-lane items on the same jobs: 3D and
    2D RK4 jobs with curvature, an expression and positions saved every step, through window
    edges, in two consecutive launches -- every output and the device work counter equal."""
    src3, gs3 = synthetic.abc_flow(24, 4)
    g2, gs2 = synthetic.double_gyre(96, 48, 4)
    ctx = capi.Context(0)
    f3 = capi.Field(ctx, 24, 24, 24, 3, dtype, nslots=4)
    f2 = capi.Field(ctx, 96, 48, 1, 2, dtype, nslots=4, address=1)
    for s in range(4):
        f3.upload(s, [src3[s, c] for c in range(3)])
        f2.upload(s, [g2[s, c] for c in range(2)])
    s3 = synthetic.uniform_seeds(31, 3000, [0.0] * 3, [2 * math.pi] * 3)
    s2 = synthetic.uniform_seeds(32, 1999, [0.02, 0.02], [1.98, 0.98])
    d3 = make_desc(3, 1, gs3, src3.shape, 0.01, 3, ("sqrt(u*u+v*v+w*w)*2-x",), save_interval=1)
    d2 = make_desc(2, 1, gs2, g2.shape, 0.01, 3, ("u*v+y",), save_interval=1)
    d3.kahan = d2.kahan = kahan

    def run(quad):
        jobs = [capi.Job(ctx, d3, s3, save_capacity=260), capi.Job(ctx, d2, s2, save_capacity=260)]
        for j in jobs:
            if not quad:
                j.set_option(capi.OPT_QUAD, 0)
        ctx.reorder(jobs)
        ctx.take_work()
        ends, _ = ctx.batch_run([(jobs[0], f3, dict(max_steps=130)), (jobs[1], f2, dict(max_steps=130))],
                                ends=True)
        nq = [ctx.get_option(capi.CTX_BATCH_QUAD_ITEMS)]
        spans = [(j, f, dict(first_window=e["window"], max_steps=130, time_start=e["time"],
                             time_start64=e["time64"], counter_start=e["counter"]))
                 for j, f, e in zip(jobs, (f3, f2), ends)]
        steps, _ = ctx.batch_run(spans)
        nq.append(ctx.get_option(capi.CTX_BATCH_QUAD_ITEMS))
        work = ctx.take_work()
        out = []
        for j in jobs:
            out += [j.read(capi.OUT_POSITIONS), j.read(capi.OUT_STATUS), j.read(capi.OUT_ABORT_STEP),
                    j.read(capi.OUT_TOTAL_CURVATURE), j.read_saved(capi.OUT_POSITIONS, 0, 260),
                    j.read_saved(capi.OUT_CURVATURE, 0, 260), j.read_saved(capi.OUT_EXPR0, 0, 260)]
            j.close()
        return [e["steps"] for e in ends] + steps, nq, work, out

    steps_q, nq_q, work_q, got = run(True)
    steps_1, nq_1, work_1, ref = run(False)
    assert steps_q == steps_1 == [130, 130, 130, 130]
    assert nq_q[0] == math.ceil(3000 / 16) + math.ceil(1999 / 16) and nq_q[1] > 0 and nq_1 == [0, 0]
    assert work_q == work_1 > 0
    for a, b in zip(ref, got):
        np.testing.assert_array_equal(a, b)
    assert (ref[1] != 0).sum() > 50  # particles stopped on the way (3D border)
    f3.close()
    f2.close()
    ctx.close()


@pytest.mark.parametrize("dtype,kahan,curv", [(capi.F16, 0, 3), (capi.F32, 1, 3), (capi.BF16, 1, 0), (capi.F32, 0, 0)])
def test_batch_per_variant_kernel_equals_kernel_of_every_variant(gpu_available, dtype, kahan, curv):
    """Spans of one specialised variant run that variant's own kernel (GR_CTX_BATCH_UNIFORM = 1)
    or the kernel of every variant (0): every output, saved column and the work counter equal,
    in wide launches (one lane per particle) and in launches of few live particles (quad
    items); a launch that mixes variants runs the kernel of every variant."""
    src, gs = synthetic.abc_flow(24, 4)
    ctx = capi.Context(0)
    f = capi.Field(ctx, 24, 24, 24, 3, dtype, nslots=4)
    for s in range(4):
        f.upload(s, [src[s, c] for c in range(3)])
    seeds = synthetic.uniform_seeds(41, 60000, [0.0] * 3, [2 * math.pi] * 3)
    d = make_desc(3, 1, gs, src.shape, 0.01, curv, ("sqrt(u*u+v*v+w*w)*2-x",), save_interval=1)
    d.kahan = kahan

    def run(uniform):
        ctx.set_option(capi.CTX_BATCH_UNIFORM, uniform)
        jobs = [capi.Job(ctx, d, seeds[:40000], save_capacity=250), capi.Job(ctx, d, seeds[40000:], save_capacity=250)]
        ctx.reorder(jobs)
        ctx.take_work()
        kernels, quads, counter, t, t64, win = [], [], 0, 0.0, 0.0, 0
        out = []
        for _ in range(5):
            ends, _ = ctx.batch_run([(j, f, dict(first_window=win, max_steps=50, time_start=t, time_start64=t64,
                                                 counter_start=counter)) for j in jobs], ends=True)
            kernels.append(ctx.get_option(capi.CTX_BATCH_LAST_KERNEL))
            quads.append(ctx.get_option(capi.CTX_BATCH_QUAD_ITEMS))
            for j in jobs:  # this launch's columns, in the order of its particles
                out += [j.read_saved(capi.OUT_POSITIONS, counter, 50), j.read_saved(capi.OUT_EXPR0, counter, 50)]
                if curv:
                    out.append(j.read_saved(capi.OUT_CURVATURE, counter, 50))
            e = ends[0]
            counter, t, t64, win = e["counter"], e["time"], e["time64"], e["window"]
            ctx.reorder(jobs)  # shrinks the live ranges: the last launches have few particles
        work = ctx.take_work()
        status = [j.read(capi.OUT_STATUS) for j in jobs]
        for j in jobs:
            out += [j.read(capi.OUT_POSITIONS), j.read(capi.OUT_ABORT_STEP)]
            if curv:
                out.append(j.read(capi.OUT_TOTAL_CURVATURE))
            j.close()
        return kernels, quads, work, counter, out + status

    k1, q1, w1, c1, got = run(1)
    k0, q0, w0, c0, ref = run(0)
    assert c0 == c1 == 250
    assert k1 == [4] * 5 and k0 == [3] * 5
    assert q0 == q1 and q1[0] == 0 and q1[-1] > 0  # wide first launch, quad items at the end
    assert w0 == w1 > 0
    for a, b in zip(ref, got):
        np.testing.assert_array_equal(a, b)
    assert sum(int((st != 0).sum()) for st in ref[-2:]) > 10000  # many particles left the domain
    # a launch mixing two variants (with and without Kahan) runs the kernel of every variant
    ctx.set_option(capi.CTX_BATCH_UNIFORM, 1)
    d2 = make_desc(3, 1, gs, src.shape, 0.01, curv, ("sqrt(u*u+v*v+w*w)*2-x",), save_interval=1)
    d2.kahan = 1 - kahan
    a, b = capi.Job(ctx, d, seeds[:1000], save_capacity=10), capi.Job(ctx, d2, seeds[:1000], save_capacity=10)
    ctx.batch_run([(a, f, dict(max_steps=10)), (b, f, dict(max_steps=10))])
    assert ctx.get_option(capi.CTX_BATCH_LAST_KERNEL) == 3
    a.close()
    b.close()
    f.close()
    ctx.close()


@pytest.mark.parametrize("dtype,integ", [(capi.F32, 1), (capi.F16, 0), (capi.BF16, 1)])
def test_batch_per_variant_kernel_2d(gpu_available, dtype, integ):
    """The 2D per-variant kernels (double gyre, clamp addressing): two jobs of one variant run
    their own kernel or the kernel of every variant with equal results, in a wide launch and in
    a launch of few live particles (quad items), and equal the FP32 oracle."""
    g2, gs2 = synthetic.double_gyre(96, 48, 4)
    ctx = capi.Context(0)
    f = capi.Field(ctx, 96, 48, 1, 2, dtype, nslots=4, address=1)
    for s in range(4):
        f.upload(s, [g2[s, c] for c in range(2)])
    seeds = synthetic.uniform_seeds(61, 50000, [0.02, 0.02], [1.98, 0.98])
    d = make_desc(2, integ, gs2, g2.shape, 0.01)

    def run(uniform, quad):
        ctx.set_option(capi.CTX_BATCH_UNIFORM, uniform)
        jobs = [capi.Job(ctx, d, seeds[:30000]), capi.Job(ctx, d, seeds[30000:])]
        for j in jobs:
            j.set_option(capi.OPT_QUAD, (1 << 29) - 1 if quad else 0)
        ctx.batch_run([(j, f, dict(max_steps=250)) for j in jobs])
        kernel, nq = ctx.get_option(capi.CTX_BATCH_LAST_KERNEL), ctx.get_option(capi.CTX_BATCH_QUAD_ITEMS)
        out = [x for j in jobs for x in (j.read(capi.OUT_POSITIONS), j.read(capi.OUT_STATUS), j.read(capi.OUT_ABORT_STEP))]
        for j in jobs:
            j.close()
        return kernel, nq, out

    ref = None
    for uniform in (1, 0):
        for quad in (False, True):
            kernel, nq, out = run(uniform, quad)
            assert kernel == (4 if uniform else 3)
            assert (nq > 0) == quad
            if ref is None:
                ref = out
            for a, b in zip(ref, out):
                np.testing.assert_array_equal(a, b)
    st = oracle_settings(2, integ, gs2, 0.01, 2.495)
    want = oracle.run(st, ROUND[dtype](g2), seeds, real64=False)
    assert want["steps"] == 250
    np.testing.assert_array_equal(np.concatenate([ref[0], ref[3]]), want["trajectories"][:, -1].astype(np.float32))
    np.testing.assert_array_equal(np.concatenate([ref[1], ref[4]]), want["status"])
    f.close()
    ctx.close()
