"""Serving: continuous batching gives every request the same result alone and in the mix,
and matches the oracle; admission respects the HBM budget."""
import math

import numpy as np
import pytest

import oracle
from pygranite_amd import capi, serving, synthetic

pytestmark = pytest.mark.gpu


def make_datasets(ctx, n=24, S=6, storage=capi.F16):
    out, raw = {}, {}
    for i in range(2):
        wind, gs = synthetic.abc_flow(n, S)
        wind = wind * (1.0 + 0.25 * i)  # two different datasets
        f = capi.Field(ctx, n, n, n, 3, storage, nslots=S)
        for s in range(S):
            f.upload(s, [wind[s, c] for c in range(3)])
        dom = [float(np.float32(n) * np.float32(g)) for g in gs]
        out[f"ds{i}"] = serving.Dataset(f, gs, dom, 1.0, S)
        raw[f"ds{i}"] = (wind, gs)
    return out, raw


def requests(seed0=0, count=8):
    reqs = []
    for i in range(count):
        n = [300, 5000, 64, 20000, 1200, 777, 9000, 128][i % 8]
        seeds = synthetic.uniform_seeds(seed0 + i, n, [0.05] * 3, [2 * math.pi - 0.05] * 3)
        reqs.append(serving.Request(rid=i, arrival=0.01 * i, dataset=f"ds{i % 2}", seeds=seeds,
                                    steps=[150, 420, 90, 700, 333, 260, 500, 610][i % 8],
                                    integrator=1 if i % 4 else 0, dt=0.005,
                                    exprs=("sqrt(u*u+v*v+w*w)",) if i % 3 == 0 else ()))
    return reqs


def test_alone_equals_in_the_mix(gpu_available):
    ctx = capi.Context(0)
    ds, raw = make_datasets(ctx)
    # the mix: all requests with overlapping arrivals and small rounds
    mix = serving.Server(ctx, ds, quantum=64)
    done_mix = {r.rid: r for r in mix.run(requests())}
    assert len(done_mix) == 8
    rep = mix.report()
    assert rep["rounds"] > 3 and rep["particle_steps"] > 0
    # each request alone, with a different quantum
    for r in requests():
        alone = serving.Server(ctx, ds, quantum=1000).run([r])[0]
        m = done_mix[r.rid]
        assert alone.steps_done == m.steps_done == r.steps
        np.testing.assert_array_equal(alone.positions, m.positions)
        np.testing.assert_array_equal(alone.status, m.status)
    # and one request against the FP32 oracle on the same FP16-rounded slices
    r = requests()[3]
    wind, gs = raw[r.dataset]
    st = dict(Space=0, Integrator=r.integrator, DeltaT=r.dt, DataTimeDistance=1.0, GridScale=gs,
              Offset=[0, 0, 0], MaximumSimulationTime=(r.steps - 0.5) * r.dt, SaveInterval=0)
    # the oracle steps while counter < MaximumSimulationTime / dt (FP32): half a step below
    # steps * dt it runs exactly r.steps steps
    ref = oracle.run(st, synthetic.round_fp16(wind), r.seeds, real64=False)
    assert ref["steps"] == r.steps
    np.testing.assert_array_equal(done_mix[r.rid].positions,
                                  ref["trajectories"][:, -1].astype(np.float32))
    np.testing.assert_array_equal(done_mix[r.rid].status, ref["status"])
    ctx.close()


def test_admission_respects_budget(gpu_available):
    ctx = capi.Context(0)
    ds, _ = make_datasets(ctx)
    reqs = requests()
    probe = serving.Server(ctx, ds)
    biggest = max(serving.job_bytes(probe._desc(r), r.seeds.shape[0], r.seeds.shape[0] >= 4096) for r in reqs)
    srv = serving.Server(ctx, ds, hbm_budget=biggest + 1, quantum=128)
    peak = [0]
    orig_round = srv.round

    def watched():
        peak[0] = max(peak[0], srv.in_use)
        orig_round()

    srv.round = watched
    srv.run(reqs)
    assert len(srv.done) == len(reqs)
    assert peak[0] <= biggest + 1
    ctx.close()



def test_job_whose_particles_all_stopped_retires_early(gpu_available):
    """A large request whose particles all leave the domain early (a uniform flow towards the
    lower x edge) is retired as soon as a regrouping has counted no alive particle (live range
    0), long before its step budget, with the result of a server that never regroups: the
    remaining steps cannot change it."""
    ctx = capi.Context(0)
    n, S = 24, 6
    wind = np.zeros((S, 3, n, n, n), np.float32)
    wind[:, 0] = -1.0
    gs = [2 * math.pi / n] * 3
    f = capi.Field(ctx, n, n, n, 3, capi.F16, nslots=S)
    for s in range(S):
        f.upload(s, [wind[s, c] for c in range(3)])
    dom = [float(np.float32(n) * np.float32(g)) for g in gs]
    ds = {"flow": serving.Dataset(f, gs, dom, 1.0, S)}
    count = 8192
    seeds = synthetic.uniform_seeds(77, count, [0.05, 0.1, 0.1], [1.0, 6.0, 6.0])

    def req():
        return serving.Request(rid=0, arrival=0.0, dataset="flow", seeds=seeds, steps=1000,
                               integrator=1, dt=0.005)

    lazy = serving.Server(ctx, ds, quantum=50, reorder=False).run([req()])[0]
    srv = serving.Server(ctx, ds, quantum=50, reorder=True, reorder_every=1)
    early = srv.run([req()])[0]
    assert (lazy.status != 0).all()  # every particle left within 1.0 / 0.005 = 200 steps
    assert early.steps_done == lazy.steps_done == 1000
    np.testing.assert_array_equal(early.positions, lazy.positions)
    np.testing.assert_array_equal(early.status, lazy.status)
    assert srv.early == 1 and srv.rounds < 1000 // 50
    ctx.close()


def test_job_bytes_matches_device_memory(gpu_available):
    """gr_job_bytes (the library's count that admission uses) agrees with the device memory a
    job's creation takes, measured with hipMemGetInfo around gr_job_create, within the
    allocator's rounding."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()

    def free_bytes():
        capi.lib()  # the engine's runtime is loaded
        assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
        return free.value

    ctx = capi.Context(0)
    alive = []  # no job is freed before the last measurement: the allocator would reuse its memory
    for curv, exprs, cap in ((3, ("sqrt(u*u+v*v+w*w)",), 8), (0, (), 1)):
        d = capi.JobDesc()
        d.dim, d.integrator, d.curvature, d.windfield_mode, d.dt, d.dtd = 3, 1, curv, 1, 0.01, 1.0
        for k in range(3):
            d.grid_scale[k], d.domain[k] = 1.0, 64.0
        d.save_interval = 1
        d.nprog = len(exprs)
        for i, e in enumerate(exprs):
            d.prog[i] = capi.compile_expr(e, dim=3, curvature=curv)
        n = 1 << 20
        want = capi.job_bytes(d, n, cap, reorder=False)
        ctx.synchronize()
        before = free_bytes()
        job = capi.Job(ctx, d, np.ones((n, 3), np.float32), save_capacity=cap)
        ctx.synchronize()
        used = before - free_bytes()
        alive.append(job)
        # per array at most a 2 MiB allocation granule of rounding
        assert want <= used <= want + 24 * (2 << 20), (want, used)
        assert capi.job_bytes(d, n, cap, reorder=True) > want
    for job in alive:
        job.close()
    ctx.close()


def test_serving_on_atmosphere_datasets(gpu_available):
    """The serving configuration's kind of dataset: atmosphere-like FP16 fields on a stretched z
    axis (48x40x20, hourly slices, dt 60 s), seeds uniform in x-y and log-uniform in height.
    Every request gives the same result alone and in the mix (jobs cross window edges), and one
    of them equals the FP32 oracle on the same FP16-rounded slices and levels."""
    nx, ny, nz, S = 48, 40, 20, 6
    ctx = capi.Context(0)
    ds, raw = {}, {}
    for k in range(2):
        wind, gs, zlev = synthetic.atmosphere(nx, ny, nz, S, seed=11 + k, dx=3000.0, dz0=20.0, ratio=1.08)
        f = capi.Field(ctx, nx, ny, nz, 3, capi.F16, nslots=S)
        f.set_z_levels(zlev)
        for s in range(S):
            f.upload(s, [wind[s, c] for c in range(3)])
        dom = [float(np.float32(nx) * np.float32(gs[0])), float(np.float32(ny) * np.float32(gs[1])), float(zlev[-1])]
        ds[f"a{k}"] = serving.Dataset(f, list(gs), dom, 3600.0, S, zlev)
        raw[f"a{k}"] = (wind, gs, zlev, dom)

    def reqs():
        out = []
        for i in range(6):
            wind, gs, zlev, dom = raw[f"a{i % 2}"]
            n = [500, 6000, 90, 15000, 2500, 4096][i]
            xy = synthetic.uniform_seeds(40 + i, n, [0.0, 0.0], [dom[0], dom[1]])
            z = synthetic.loguniform_heights(50 + i, n, float(zlev[1]), float(zlev[-1]))
            seeds = np.ascontiguousarray(np.concatenate([xy, z[:, None]], axis=1), np.float32)
            out.append(serving.Request(rid=i, arrival=0.005 * i, dataset=f"a{i % 2}", seeds=seeds,
                                       steps=[100, 130, 60, 120, 80, 110][i],
                                       integrator=0 if i == 2 else 1, dt=60.0))
        return out

    mix = {r.rid: r for r in serving.Server(ctx, ds, quantum=50).run(reqs())}
    assert len(mix) == 6
    for r in reqs():
        alone = serving.Server(ctx, ds, quantum=1000).run([r])[0]
        assert alone.steps_done == mix[r.rid].steps_done == r.steps
        np.testing.assert_array_equal(alone.positions, mix[r.rid].positions)
        np.testing.assert_array_equal(alone.status, mix[r.rid].status)
    r = reqs()[3]
    wind, gs, zlev, dom = raw[r.dataset]
    st = dict(Space=0, Integrator=r.integrator, DeltaT=r.dt, DataTimeDistance=3600.0, GridScale=list(gs),
              Offset=[0, 0, 0], MaximumSimulationTime=(r.steps - 0.5) * r.dt, SaveInterval=0)
    ref = oracle.run(st, synthetic.round_fp16(wind), r.seeds, real64=False, zlevels=zlev)
    assert ref["steps"] == r.steps
    np.testing.assert_array_equal(mix[r.rid].positions, ref["trajectories"][:, -1].astype(np.float32))
    np.testing.assert_array_equal(mix[r.rid].status, ref["status"])
    assert (ref["status"] != 0).sum() > 10 and (ref["status"] == 0).sum() > 10  # stops and survivors
    ctx.close()
