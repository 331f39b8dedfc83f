#!/usr/bin/env python3
"""Benchmark: RK4 3D particle-steps/s on the time-modulated ABC flow (one MI355X).

Workload (the 3D ABC configuration): 128^3 grid over [0, 2*pi)^3, 16 slices one time unit
apart, 1,048,576 uniform seeds (SplitMix64 seed 1), RK4 with dt = 0.005, 2000 steps, per-step
curvature (IndividualAndTotalCurvature) and the expression sqrt(u*u+v*v+w*w), saved at every
step into a device-resident column ring. All slices are resident in HBM before the timed region.

A bench "step" is one launch advancing every particle by ceil(2000 / --steps) integration
steps, so the timed --steps launches are always the whole 2000-step job from its first step
(the last launch is shorter when the division is not exact). The warm-up launches run on a
separate copy of the job. The timed job is repeated --repeats times from fresh state; `value`
is the median.

Every seed of this workload leaves the domain (3D border aborts) within about 1000 steps, so
the headline counts integrated particle-steps: steps a particle began alive, counted on the
device by the step kernels. The nominal count (particles x steps) is reported beside it.

Forms, all on identical inputs:
  plain_fp32    FP32 slices sampled slot by slot, particles in caller order, one job per launch
                (the plain FP32 GPU form: the reference's arithmetic and layout)
  fp32_local    FP32 slices, time-paired layout, Morton-ordered particles (bit for bit plain_fp32)
  fp32k         fp32_local with Kahan-compensated positions
  fp16          FP16 slices, time-paired layout, Morton-ordered particles
  fp16_batched  the fp16 form run by the persistent batched kernel, the particle set split into
                16 jobs of 65,536
  fp32k_batched the fp32k form run by the batched kernel, 16 jobs
  fp32_tile     fp32_local with the per-workgroup LDS field tile (GR_OPT_TILE, --tile-bytes)
  fp16_tile     fp16 with the LDS field tile
Final positions of every form are compared with the FP64 CPU oracle on 16,384 particles:
domain-relative error over particles whose termination agrees, and the termination mismatches
classified from both step-by-step paths (divergence, threshold, other). The value is the fastest
form other than plain_fp32 whose RMS error meets 1e-4 with no unexplained ("other") mismatch;
every form's rate is reported beside it.

Also reported: the 2D double-gyre configuration on the GPU and as the CPU baseline in full
(65,536 particles x 1000 Euler steps, all host threads and one), the CPU baseline on a bounded
sample of the headline workload, lite batched-jobs and serving runs, and the serving
configuration (Poisson arrivals at 50 jobs/s for --serve-default-duration seconds onto four
atmosphere datasets; p50/p99 job latency) in a child process.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
ERR_TARGET = 1e-4      # domain-relative RMS position error bound against the FP64 oracle
FORMS = ("plain_fp32", "fp32_local", "fp32k", "fp16", "fp16_batched", "fp32k_batched", "fp32_tile", "fp16_tile")
KAHAN_FORMS = ("fp32k", "fp32k_batched")
# VALU issue ceiling: 256 CUs x 2.4 GHz x one wave64 (non-packed FP32) instruction per clock per CU
VALU_PEAK_WAVE_INSTS = 256 * 2.4e9
EXPRS = ["sqrt(u*u+v*v+w*w)"]
DT = np.float32(0.005)


def build_desc(capi, gs, n, save_interval=1, curvature=3, exprs=EXPRS, kahan=False):
    d = capi.JobDesc()
    d.dim, d.integrator, d.curvature, d.windfield_mode = 3, 1, curvature, 1
    d.dt, d.dtd = float(DT), 1.0
    d.kahan = 1 if kahan else 0
    for k in range(3):
        d.grid_scale[k] = gs[k]
        d.domain[k] = float(np.float32(n) * np.float32(gs[k]))
    d.save_interval = save_interval
    d.nprog = len(exprs)
    for i, e in enumerate(exprs):
        d.prog[i] = capi.compile_expr(e, dim=3, curvature=curvature)
    return d


def algorithmic_bytes(storage_bytes: int, outputs: int) -> int:
    """Bytes per RK4-3D particle-step: 4 stages x 2 slices x 3 comps x 8 corners x s
    + position read/write and status (2*3*4 + 4) + saved outputs."""
    return 4 * 2 * 3 * 8 * storage_bytes + (2 * 3 * 4 + 4) + outputs


class Schedule:
    """The reference's FP32 host schedule (window edges at Time >= max_windfield_time)."""

    def __init__(self):
        self.t = np.float32(0)
        self.counter = 0
        self.window = 0
        self.mwt = np.float32(1.0)

    def chunk(self, k_max):
        """Splits up to k_max steps at the next window edge: returns (k, edge_reached)."""
        k, t = 0, self.t
        while k < k_max:
            k += 1
            t = np.float32(t + DT)
            if t >= self.mwt:
                return k, True, t
        return k, False, t

    def advance(self, k, edge, t):
        self.counter += k
        self.t = t
        if edge:
            self.window += 1
            self.mwt = np.float32(self.mwt + np.float32(1.0))


def make_field(capi, ctx, wind, dtype):
    S, n = wind.shape[0], wind.shape[2]
    f = capi.Field(ctx, n, n, n, 3, dtype=dtype, nslots=S, address=0)
    for s in range(S):
        f.upload(s, [wind[s, c] for c in range(3)])
    for s in range(S):
        f.wait(s)
    return f


def run_single(capi, ctx, f, seeds, gs, n, chunk, steps, warmup, paired, reorder, reorder_every=0,
               regroup_min=0, total=2000, repeats=1, cu_waves=0, reuse=0, kahan=False, quad=-1, tile=0,
               tile_halo=1, edge_min=0, block=0):
    def new_job():
        j = capi.Job(ctx, build_desc(capi, gs, n, kahan=kahan), seeds, save_capacity=chunk)
        j.set_option(capi.OPT_PAIRED, 1 if paired else 0)
        if block:
            j.set_launch(block)
        if quad >= 0:
            j.set_option(capi.OPT_QUAD, quad)
        if tile:
            j.set_option(capi.OPT_TILE, tile)
            j.set_option(capi.OPT_TILE_HALO, tile_halo)
        if cu_waves:
            j.set_option(capi.OPT_CU_WAVES, cu_waves)
        if reuse > 0:
            j.set_option(capi.OPT_REUSE, 1)
        if reorder:
            j.reorder()
        return j

    st = {}

    def reset():
        st.update(job=new_job(), sch=Schedule(), count=0, large=True)

    def launch():
        job, sch = st["job"], st["sch"]
        left = min(chunk, total - sch.counter)
        regrouped = False
        while left > 0:
            k, edge, t = sch.chunk(left)
            w = capi.Window()
            w.wind, w.wind_back, w.wind_front = f.h, sch.window, sch.window + 1
            w.time_back, w.time_front = float(np.float32(sch.window)), float(np.float32(sch.window + 1))
            w.time_back64, w.time_front64 = float(sch.window), float(sch.window + 1)
            w.time_start = w.time_start64 = float(sch.t)
            w.counter_start, w.nsteps = sch.counter, k
            job.advance(w)
            sch.advance(k, edge, t)
            left -= k
            if edge and reorder and job.live_range() >= edge_min:
                job.reorder()  # regroup by cell at window edges (inside the timing) while at
                regrouped = True  # least edge_min particles are live
        st["count"] += 1
        if reorder and reorder_every and not regrouped and st["count"] % reorder_every == 0 and st["large"]:
            if job.live_range() >= regroup_min:
                job.reorder()  # and every reorder_every launches while the live range is large
            else:
                st["large"] = False

    reset()
    for _ in range(warmup):
        launch()
    st["job"].close()
    runs = []
    for rep in range(repeats):
        # each timed job starts from step 0 on fresh state
        reset()
        ctx.synchronize()
        ctx.take_work()
        ctx.take_kernel_ms()
        t0 = time.perf_counter()
        for _ in range(steps):
            launch()
        ctx.synchronize()
        wall = time.perf_counter() - t0
        work = ctx.take_work()
        kernel_ms, launches = ctx.take_kernel_ms()  # HIP events around every step-kernel launch
        r = {"wall_s": wall, "kernel_ms": kernel_ms, "launches": launches,
             "steps_done": st["sch"].counter, "work": work}
        if tile:
            tiled, untiled = ctx.take_tile_stats()
            r["tile_workgroup_steps"] = {"with_tile": tiled, "without_tile": untiled}
        if rep == repeats - 1:
            job = st["job"]
            r.update(positions=job.read(capi.OUT_POSITIONS), status=job.read(capi.OUT_STATUS),
                     abort_step=job.read(capi.OUT_ABORT_STEP))
        st["job"].close()
        runs.append(r)
    return median_run(runs)


def median_run(runs):
    """The run of median wall time, with every run's wall time and the final state of the last
    run (identical in every run: results do not depend on timing)."""
    order = sorted(range(len(runs)), key=lambda i: runs[i]["wall_s"])
    out = dict(runs[order[len(runs) // 2]])
    last = runs[-1]
    for k in ("positions", "status", "abort_step"):
        if k in last:
            out[k] = last[k]
    out["repeat_walls_s"] = [r["wall_s"] for r in runs]
    return out


def run_batched(capi, ctx, f, seeds, gs, n, chunk, steps, warmup, njobs, reorder_every=0, regroup_min=0,
                total=2000, repeats=1, kahan=False, edge_min=0):
    parts = np.array_split(np.arange(seeds.shape[0]), njobs)

    def new_jobs():
        out = [capi.Job(ctx, build_desc(capi, gs, n, kahan=kahan), seeds[idx], save_capacity=chunk)
               for idx in parts]
        ctx.reorder(out)
        return out

    state = {}

    def reset():
        state.update(jobs=new_jobs(), window=0, time=0.0, time64=0.0, counter=0, launches=0, large=True)

    def launch():
        jobs = state["jobs"]
        k = min(chunk, total - state["counter"])
        if k <= 0:
            return
        spans = [(j, f, dict(first_window=state["window"], max_steps=k,
                             time_start=state["time"], time_start64=state["time64"],
                             counter_start=state["counter"])) for j in jobs]
        ends, _ = ctx.batch_run(spans, timed=False, ends=True)
        e = ends[0]
        state["launches"] += 1
        # live ranges only shrink: once below regroup_min they are not queried again
        if (reorder_every and state["large"] and state["launches"] % reorder_every == 0
                and sum(j.live_range() for j in jobs) < regroup_min):
            state["large"] = False
        edge = e["window"] != state["window"] and sum(j.live_range() for j in jobs) >= edge_min
        if edge or (reorder_every and state["large"] and state["launches"] % reorder_every == 0):
            ctx.reorder(jobs)  # every job regrouped by cell in one pass: at window edges, and
            # every reorder_every launches while the live range is large
        state.update(window=e["window"], time=e["time"], time64=e["time64"], counter=e["counter"])

    reset()
    for _ in range(warmup):
        launch()
    for j in state["jobs"]:
        j.close()
    runs = []
    for rep in range(repeats):
        reset()
        ctx.synchronize()
        ctx.take_work()
        ctx.take_kernel_ms()
        t0 = time.perf_counter()
        for _ in range(steps):
            launch()
        ctx.synchronize()
        wall = time.perf_counter() - t0
        work = ctx.take_work()
        kernel_ms, launches = ctx.take_kernel_ms()
        r = {"wall_s": wall, "kernel_ms": kernel_ms, "launches": launches,
             "steps_done": state["counter"], "work": work}
        jobs = state["jobs"]
        if rep == repeats - 1:
            r.update(positions=np.concatenate([j.read(capi.OUT_POSITIONS) for j in jobs]),
                     status=np.concatenate([j.read(capi.OUT_STATUS) for j in jobs]),
                     abort_step=np.concatenate([j.read(capi.OUT_ABORT_STEP) for j in jobs]))
        for j in jobs:
            j.close()
        runs.append(r)
    return median_run(runs)


def fp64_error(wind, seeds_sub, gs, steps_done, forms, storage_round=None, kahan_forms=()):
    """Final-position error of GPU forms against the FP64 oracle on the same particles.

    forms: {name: result dict with positions/status/abort_step of at least len(seeds_sub)
    particles}; storage_round: {name: function rounding the field as that form stores it};
    kahan_forms: names of the forms with Kahan-compensated positions. Returns {name: error
    summary} (termination_and_error)."""
    import oracle
    # the oracle steps while counter < MaximumSimulationTime / dt (FP32): half a step below
    # steps_done * dt it runs exactly steps_done steps
    st = dict(Space=0, Integrator=1, DeltaT=float(DT), DataTimeDistance=1.0, GridScale=gs,
              Offset=[0, 0, 0], MaximumSimulationTime=(steps_done - 0.5) * float(DT),
              SaveInterval=0)
    threads = min(os.cpu_count() or 1, 16)
    sub = seeds_sub.shape[0]
    ref = oracle.run(st, wind, seeds_sub, real64=True, nthreads=threads)
    if ref["steps"] != steps_done:
        return {"error": f"oracle ran {ref['steps']} steps, GPU {steps_done}"}
    out = {}
    for name, res in forms.items():
        rnd = (storage_round or {}).get(name)
        field = rnd(wind) if rnd else wind

        def paths(idx, field=field, kahan=name in kahan_forms):
            # every step of the mismatched particles: the FP64 oracle on the exact field, and the
            # FP32 restatement of the GPU form (bit for bit the GPU path: the field as the form
            # stores it, the form's summation)
            s1 = dict(st, SaveInterval=1)
            r64 = oracle.run(s1, wind, seeds_sub[idx], real64=True, nthreads=threads)
            r32 = oracle.run(s1, field, seeds_sub[idx], real64=False, kahan=kahan, nthreads=threads)
            return r64, r32

        out[name] = termination_and_error(
            res["positions"][:sub], res["status"][:sub], res["abort_step"][:sub],
            ref["trajectories"][:, -1, :], ref["status"], ref["abort_step"], [2 * math.pi] * 3,
            paths, {"steps": int(steps_done)})
    return out


def termination_and_error(pos, status, ab, pos_ref, status_ref, ab_ref, extent, paths, extra=None):
    """Domain-relative max / RMS position error over particles whose status and abort step
    agree with the FP64 oracle, and a classification of the termination mismatches, from both
    paths of each mismatched particle up to the first of the two stops (step a):
      divergence  the paths are more than one step length apart at step a: the stops differ
                  because the trajectories had already separated (chaotic growth of rounding or
                  storage differences);
      threshold   within one step length and the same abort reason: which step crosses the
                  domain edge (or stalls) is decided by less than the paths' difference;
      other       the reasons differ while the paths are within a step -- a genuine
                  disagreement."""
    ext = np.asarray(extent, np.float64)
    same = (status_ref == status) & (ab_ref == ab)
    d = np.abs(np.asarray(pos, np.float64)[same] - np.asarray(pos_ref, np.float64)[same]) / ext
    mism = np.nonzero(~same)[0]
    cls = {"divergence": 0, "threshold": 0, "other": 0}
    gap = []
    if mism.size:
        r64, r32 = paths(mism)
        T64, T32 = r64["trajectories"], r32["trajectories"]  # (m, steps + 1, 3), column k = after k steps
        for m in range(mism.size):
            a = [x for x in (int(r64["abort_step"][m]), int(r32["abort_step"][m])) if x >= 0]
            k = min(a) if a else T64.shape[1] - 1  # position column of the first stop
            k = max(1, min(k, T64.shape[1] - 1))
            sep = np.max(np.abs(T64[m, k] - T32[m, k]))
            step_len = np.max(np.abs(T64[m, k] - T64[m, k - 1]))
            gap.append(sep / max(step_len, 1e-30))
            if sep > step_len:
                cls["divergence"] += 1
            elif r64["status"][m] == r32["status"][m] or 0 in (r64["status"][m], r32["status"][m]):
                cls["threshold"] += 1
            else:
                cls["other"] += 1
    out = {"rms": float(np.sqrt(np.mean(d * d))) if d.size else None,
           "max": float(d.max()) if d.size else None,
           "particles": int(len(status)),
           "termination_agree": int(same.sum()),
           "termination_mismatch": int(mism.size),
           "mismatch_divergence": cls["divergence"],
           "mismatch_threshold": cls["threshold"],
           "mismatch_other": cls["other"],
           "median_separation_in_steps": float(np.median(gap)) if gap else None,
           "alive_at_end": int((status_ref == 0).sum())}
    if extra:
        out.update(extra)
    return out


def cpu_baseline(wind, gs):
    """Plain FP32 CPU path: the oracle's FP32 restatement on the host threads over a bounded
    sample of the same workload (131,072 particles x 1,000 steps)."""
    import oracle
    from pygranite_amd import synthetic
    n, steps = 131072, 1000
    seeds = synthetic.uniform_seeds(11, n, [0.0] * 3, [2 * math.pi] * 3)
    st = dict(Space=0, Integrator=1, DeltaT=float(DT), DataTimeDistance=1.0, GridScale=gs,
              Offset=[0, 0, 0], MaximumSimulationTime=steps * float(DT), SaveInterval=1,
              CurvatureMode=3, AdditionalCompute=EXPRS)
    threads = min(os.cpu_count() or 1, 16)
    t0 = time.perf_counter()
    r = oracle.run(st, wind[:7], seeds, real64=False, nthreads=threads)
    el = time.perf_counter() - t0
    # integrated particle-steps, as on the GPU: steps begun alive, the aborting step included
    def integrated(res):
        alive = res["status"] == 0
        return int(np.where(alive, res["steps"], res["abort_step"].astype(np.int64) + 1).sum())

    work = integrated(r)
    # one thread on a sixteenth of the sample (about the same wall time)
    n1 = n // 16
    t0 = time.perf_counter()
    r1 = oracle.run(st, wind[:7], seeds[:n1], real64=False, nthreads=1)
    el1 = time.perf_counter() - t0
    w1 = integrated(r1)
    return {"value": work / el, "unit": "particle-steps/s", "cores": threads,
            "kind": "port", "nominal_value": n * r["steps"] / el,
            "sample": f"{n} particles x {r['steps']} RK4 steps of the same 3D ABC workload with "
                      f"curvature + expression saved every step (FP32 oracle restatement, "
                      f"{threads} threads, {el:.1f} s; {work} integrated particle-steps)",
            "one_thread": {"value": w1 / el1, "particles": n1, "seconds": el1, "integrated_particle_steps": w1}}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def gyre_config():
    """The 2D double-gyre configuration: 512 x 256 nodes over [0, 2] x [0, 1], 11 slices one
    time unit apart, 65,536 seeds uniform in the interior (SplitMix64 seed 0), Euler, dt 0.01,
    1000 steps, Block borders, saved every 10th step."""
    from pygranite_amd import synthetic
    wind, gs = synthetic.double_gyre(512, 256, 11)
    seeds = synthetic.uniform_seeds(0, 65536, [gs[0], gs[1]], [2.0 - gs[0], 1.0 - gs[1]])
    st = dict(Space=1, Integrator=0, DeltaT=0.01, DataTimeDistance=1.0, GridScale=gs, Offset=[0, 0],
              MaximumSimulationTime=10.0, SaveInterval=10)
    return wind, gs, seeds, st


def gyre_cpu_baseline():
    """The double-gyre configuration in full on the CPU (the oracle's FP32 restatement), on the
    host threads (at most 16) and on one thread."""
    import oracle
    wind, gs, seeds, st = gyre_config()
    out = {"cpu_model": cpu_model(), "host_threads_visible": os.cpu_count()}
    for label, threads in (("all", min(os.cpu_count() or 1, 16)), ("one", 1)):
        t0 = time.perf_counter()
        r = oracle.run(st, wind, seeds, real64=False, nthreads=threads)
        el = time.perf_counter() - t0
        alive = r["status"] == 0
        work = int(np.where(alive, r["steps"], r["abort_step"].astype(np.int64) + 1).sum())
        out[label] = {"threads": threads, "seconds": el, "steps": int(r["steps"]),
                      "particle_steps_per_s": work / el, "nominal_particle_steps_per_s":
                      seeds.shape[0] * r["steps"] / el, "integrated_particle_steps": work}
        out["saved"] = r["positions"][1:].astype(np.float32)  # (columns, N, 2), start set dropped
    return out


def gyre_gpu(capi, ctx, reference=None):
    """The double-gyre configuration on the GPU: the whole 1000-step job in one launch (the
    launch crosses the ten window edges), plain FP32 (slots sampled separately, caller order)
    and FP32 with the locality layout; kernel-only time and time including the read of the
    saved columns. With `reference` (the CPU path's saved columns, (columns, N, 2)) every saved
    position is compared bit for bit."""
    wind, gs, seeds, st = gyre_config()
    f = make_field2(capi, ctx, wind, capi.F32)
    out = {}
    for form, paired in (("plain_fp32", 0), ("fp32_local", 1)):
        d = capi.JobDesc()
        d.dim, d.integrator, d.windfield_mode, d.dt, d.dtd = 2, 0, 1, 0.01, 1.0
        d.grid_scale[0], d.grid_scale[1], d.grid_scale[2] = gs[0], gs[1], gs[0]
        d.domain[0] = float(np.float32(512) * np.float32(gs[0]))
        d.domain[1] = float(np.float32(256) * np.float32(gs[1]))
        d.save_interval = 10
        job = capi.Job(ctx, d, seeds, save_capacity=101)
        job.set_option(capi.OPT_PAIRED, paired)
        if paired:
            job.reorder()
        ctx.synchronize()
        ctx.take_work()
        ctx.take_kernel_ms()
        t0 = time.perf_counter()
        ends, _ = ctx.batch_run([(job, f, dict(max_steps=1000, max_sim_time=10.0))], timed=False, ends=True)
        ctx.synchronize()
        t_kernel = time.perf_counter() - t0
        cols = job.read_saved(capi.OUT_POSITIONS, 0, job.saved_columns())
        t_e2e = time.perf_counter() - t0
        work = ctx.take_work()
        kms, _ = ctx.take_kernel_ms()
        out[form] = {"steps": ends[0]["steps"], "device_ms": kms, "wall_s": t_kernel,
                     "particle_steps_per_s": work / (kms / 1e3),
                     "end_to_end_s": t_e2e, "saved_columns": int(cols.shape[0]),
                     "end_to_end_particle_steps_per_s": work / t_e2e,
                     "bytes_per_particle_step": 84}
        if reference is not None:
            out[form]["bitexact_vs_cpu_fp32"] = bool(cols.shape == reference.shape and
                                                     np.array_equal(cols, reference))
        job.close()
    f.close()
    return out


def batched_jobs_lite(capi, ctx, seed=3, njobs=64):
    """Mixed jobs (half 2D double gyre, half 3D ABC; Euler:RK4 = 1:3; FP16; particle counts
    log-uniform [1e3, 1e5]; steps uniform [100, 1000]) in one persistent launch vs the same jobs
    as sequential plain FP32 launches."""
    from pygranite_amd import synthetic
    u = synthetic.uniform01(seed, 4 * njobs).reshape(njobs, 4)
    g2, gs2 = synthetic.double_gyre(512, 256, 11)
    g3, gs3 = synthetic.abc_flow(128, 16)
    fields = {(2, capi.F16): make_field2(capi, ctx, g2, capi.F16),
              (2, capi.F32): make_field2(capi, ctx, g2, capi.F32),
              (3, capi.F16): make_field(capi, ctx, g3, capi.F16),
              (3, capi.F32): make_field(capi, ctx, g3, capi.F32)}
    specs = []
    for i in range(njobs):
        dim = 2 if i % 2 == 0 else 3
        integ = 0 if u[i, 0] < 0.25 else 1
        n = int(round(math.exp(math.log(1e3) + u[i, 1] * (math.log(1e5) - math.log(1e3)))))
        steps = int(100 + u[i, 2] * 900)
        specs.append((dim, integ, n, steps))

    def make_jobs(dtype, reorder):
        jobs = []
        for i, (dim, integ, n, steps) in enumerate(specs):
            gs = gs2 if dim == 2 else gs3
            if dim == 2:
                seeds = synthetic.uniform_seeds(1000 + i, n, [0.01, 0.01], [1.99, 0.99])
                d = capi.JobDesc()
                d.dim, d.integrator, d.windfield_mode = 2, integ, 1
                d.dt, d.dtd = 0.01, 1.0
                d.grid_scale[0], d.grid_scale[1], d.grid_scale[2] = gs[0], gs[1], gs[0]
                d.domain[0] = float(np.float32(512) * np.float32(gs[0]))
                d.domain[1] = float(np.float32(256) * np.float32(gs[1]))
            else:
                seeds = synthetic.uniform_seeds(1000 + i, n, [0.0] * 3, [2 * math.pi] * 3)
                d = capi.JobDesc()
                d.dim, d.integrator, d.windfield_mode = 3, integ, 1
                d.dt, d.dtd = float(DT), 1.0
                for k in range(3):
                    d.grid_scale[k] = gs[k]
                    d.domain[k] = float(np.float32(128) * np.float32(gs[k]))
            j = capi.Job(ctx, d, seeds)
            if reorder and n >= 4096:
                j.reorder()
            jobs.append((j, fields[(dim, dtype)], steps, n))
        return jobs

    # batched FP16: one persistent launch
    jobs = make_jobs(capi.F16, True)
    ctx.synchronize()
    ctx.take_work()
    t0 = time.perf_counter()
    steps_run, ms = ctx.batch_run([(j, f, dict(max_steps=s)) for j, f, s, _ in jobs])
    wall_b = time.perf_counter() - t0
    work_b = ctx.take_work()
    ps = sum(k * n for k, (_, _, _, n) in zip(steps_run, jobs))
    for j, *_ in jobs:
        j.close()
    # sequential plain FP32 launches (one job per launch, slot-by-slot sampling, caller order)
    jobs = make_jobs(capi.F32, False)
    for j, *_ in jobs:
        j.set_option(capi.OPT_PAIRED, 0)
    ctx.synchronize()
    t0 = time.perf_counter()
    for (j, f, s, n) in jobs:
        ctx.batch_run([(j, f, dict(max_steps=s))], timed=False)
    ctx.synchronize()
    wall_s = time.perf_counter() - t0
    work_s = ctx.take_work()
    for j, *_ in jobs:
        j.close()
    for f in fields.values():
        f.close()
    return {"jobs": njobs, "nominal_particle_steps": ps, "integrated_particle_steps": work_b,
            "integrated_particle_steps_plain": work_s, "batched_fp16_s": wall_b,
            "sequential_plain_fp32_s": wall_s, "speedup": wall_s / wall_b,
            "batched_psps": work_b / wall_b}


def make_field2(capi, ctx, wind, dtype):
    S, ny, nx = wind.shape[0], wind.shape[2], wind.shape[3]
    f = capi.Field(ctx, nx, ny, 1, 2, dtype=dtype, nslots=S, address=1)
    for s in range(S):
        f.upload(s, [wind[s, c] for c in range(2)])
    for s in range(S):
        f.wait(s)
    return f


def serving_lite(capi, ctx, rate=50.0, duration=8.0, seed=4):
    """Poisson arrivals at `rate` jobs/s for `duration` s onto two resident 3D datasets
    (ABC 128^3 x 16 slices, FP16); job sizes log-uniform [1e3, 1e5] particles, uniform
    [100, 1000] steps, RK4:Euler = 3:1."""
    from pygranite_amd import serving, synthetic
    g3, gs3 = synthetic.abc_flow(128, 16)
    ds = {}
    for i in range(2):
        f = make_field(capi, ctx, g3 * (1.0 + 0.2 * i), capi.F16)
        ds[f"d{i}"] = serving.Dataset(f, gs3, [float(np.float32(128) * np.float32(g))
                                              for g in gs3], 1.0, 16)
    arr = serving.poisson_arrivals(seed, rate, duration)
    u = synthetic.uniform01(seed + 1, 4 * len(arr)).reshape(len(arr), 4)
    reqs = []
    for i, t in enumerate(arr):
        n = int(round(math.exp(math.log(1e3) + u[i, 0] * (math.log(1e5) - math.log(1e3)))))
        reqs.append(serving.Request(
            rid=i, arrival=float(t), dataset=f"d{int(u[i, 1] * 2)}",
            seeds=synthetic.uniform_seeds(5000 + i, n, [0.0] * 3, [2 * math.pi] * 3),
            steps=int(100 + u[i, 2] * 900), integrator=0 if u[i, 3] < 0.25 else 1,
            dt=float(DT)))
    srv = serving.Server(ctx, ds, hbm_budget=32 << 30, quantum=200)
    ctx.take_work()
    srv.run(reqs)
    rep = srv.report()
    rep["integrated_particle_steps"] = ctx.take_work()
    rep["integrated_particle_steps_per_s"] = rep["integrated_particle_steps"] / rep["wall_s"]
    rep.update(arrival_rate_per_s=rate, duration_s=duration,
               offered_particle_steps_per_s=sum(r.seeds.shape[0] * r.steps for r in reqs) / duration)
    for d in ds.values():
        d.field.close()
    return rep


def atmosphere_workload(capi, ctx, args):
    """Precision form on the atmosphere-like field: 512x512x64 (stretched z, ratio 1.05),
    25 hourly slices (24 windows = 1440 steps of 60 s), 16,777,216 seeds (uniform x-y,
    log-uniform height), RK4, Kahan positions, FP64 time, the expression
    w*1000/(abs(u)+abs(v)+1) saved every step. Storage FP32 / FP16 / BF16; error of each
    against the FP64 oracle on a subsample. Prints one JSON line."""
    import torch
    from pygranite_amd import synthetic
    nx = ny = args.atmo_nx
    nz, S = 64, 25
    dev = torch.device("cuda")
    t0 = time.perf_counter()
    wind_t, gs, zlev = synthetic.atmosphere(nx, ny, nz, S, seed=1, device=dev)
    torch.cuda.synchronize()
    print(f"# atmosphere field generated in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    dom = [float(np.float32(nx) * np.float32(gs[0])), float(np.float32(ny) * np.float32(gs[1])),
           float(zlev[-1])]
    n = args.atmo_particles
    xy = synthetic.uniform_seeds(21, n, [0.0, 0.0], [dom[0], dom[1]])
    z = synthetic.loguniform_heights(22, n, float(zlev[1]), float(zlev[-1]))
    seeds = np.ascontiguousarray(np.concatenate([xy, z[:, None]], axis=1), np.float32)
    expr = "w*1000/(abs(u)+abs(v)+1)"
    dt, chunk = np.float32(60.0), args.chunk or 120  # steps per launch (two 60-step windows)
    total_steps = 1440
    out = {"workload": f"atmosphere {nx}x{ny}x{nz} stretched z, {S} slices, {n} particles, RK4 dt 60 s, "
                       f"{total_steps} steps, Kahan + FP64 time, expression saved every step"}
    finals = {}
    for name, dtype in (("fp32", capi.F32), ("fp16", capi.F16), ("bf16", capi.BF16)):
        f = capi.Field(ctx, nx, ny, nz, 3, dtype=dtype, nslots=S, address=0)
        f.set_z_levels(zlev)
        torch.cuda.synchronize()  # the engine's copies are not ordered after torch's stream
        for s in range(S):
            f.upload_device(s, [wind_t[s, c].data_ptr() for c in range(3)])
        ctx.synchronize()
        d = capi.JobDesc()
        d.dim, d.integrator, d.windfield_mode = 3, 1, 1
        d.dt, d.dtd = float(dt), 3600.0
        d.kahan, d.time64 = 1, 1
        for k in range(3):
            d.grid_scale[k] = gs[k]
            d.domain[k] = dom[k]
        d.save_interval = 1
        d.nprog = 1
        d.prog[0] = capi.compile_expr(expr, dim=3, allow_abs=True)
        job = capi.Job(ctx, d, seeds, save_capacity=chunk)
        job.reorder()
        ctx.synchronize()
        ctx.take_work()
        ctx.take_kernel_ms()
        st = dict(window=0, time=0.0, time64=0.0, counter=0)
        t0 = time.perf_counter()
        while st["counter"] < total_steps:
            ends, _ = ctx.batch_run([(job, f, dict(first_window=st["window"],
                                                   max_steps=min(chunk, total_steps - st["counter"]),
                                                   time_start=st["time"], time_start64=st["time64"],
                                                   counter_start=st["counter"]))], ends=True, timed=False)
            e = ends[0]
            if e["window"] != st["window"]:
                job.reorder()
            st.update(window=e["window"], time=e["time"], time64=e["time64"], counter=e["counter"])
            if e["done"] or e["steps"] == 0:
                break
        ctx.synchronize()
        wall = time.perf_counter() - t0
        work = ctx.take_work()
        kms, nl = ctx.take_kernel_ms()
        out[name] = {"particle_steps_per_s": work / wall, "nominal_particle_steps_per_s": n * st["counter"] / wall,
                     "integrated_fraction": work / (n * max(1, st["counter"])), "wall_s": wall,
                     "device_s": kms / 1e3, "launches": nl, "steps": st["counter"],
                     "field_bytes_per_slice": f.bytes_per_slot()}
        print(f"#   atmosphere {name}: {out[name]['particle_steps_per_s']:.3e} integrated particle-steps/s "
              f"({out[name]['integrated_fraction']:.3f} of nominal)", file=sys.stderr, flush=True)
        finals[name] = (job.read(capi.OUT_POSITIONS), job.read(capi.OUT_STATUS),
                        job.read(capi.OUT_ABORT_STEP), st["counter"])
        job.close()
        f.close()
    if not args.no_error:
        import oracle
        sub = min(args.atmo_error_particles, n)
        wind_h = wind_t.cpu().numpy()
        del wind_t
        st = dict(Space=0, Integrator=1, DeltaT=float(dt), DataTimeDistance=3600.0,
                  GridScale=list(gs), Offset=[0, 0, 0], SaveInterval=0,
                  MaximumSimulationTime=float(total_steps) * float(dt) + 1.0)
        threads = min(os.cpu_count() or 1, 16)
        ref = oracle.run(st, wind_h, seeds[:sub], real64=True, zlevels=zlev, time64=True, nthreads=threads)

        seeds_sub = seeds[:sub]
        rounders = {"fp16": synthetic.round_fp16, "bf16": synthetic.round_bf16}
        for name, (pos, status, ab, steps) in finals.items():
            rnd = rounders.get(name)

            def paths(idx, field=rnd(wind_h) if rnd else wind_h):
                s1 = dict(st, SaveInterval=1)
                r64 = oracle.run(s1, wind_h, seeds_sub[idx], real64=True, zlevels=zlev, time64=True,
                                 nthreads=threads)
                r32 = oracle.run(s1, field, seeds_sub[idx], real64=False, kahan=True, zlevels=zlev,
                                 time64=True, nthreads=threads)
                return r64, r32

            out[name]["fp64_error"] = termination_and_error(
                pos[:sub], status[:sub], ab[:sub], ref["trajectories"][:, -1, :], ref["status"],
                ref["abort_step"], dom, paths)
    out["speedup_fp16_vs_fp32"] = out["fp16"]["particle_steps_per_s"] / out["fp32"]["particle_steps_per_s"]
    print(json.dumps(out), flush=True)


def batch_workload(capi, ctx, args):
    """The batched configuration: 256 independent jobs in one persistent launch. Particle
    counts log-uniform in [1e3, 4e6], step counts uniform in [100, 5000] (SplitMix64 seed 3);
    even jobs 2D double gyre (512x256, dt 0.01), odd jobs 3D ABC (128^3, dt 0.005); Euler:RK4 =
    1:3; FP16 storage, launch configuration from the autotuner. Both analytic fields are
    sampled over enough slices for the longest job (51 and 26). Compared with the same jobs as
    256 sequential plain FP32 launches (one job each, FP32 slots sampled separately, caller
    order). Prints one JSON line."""
    from pygranite_amd import synthetic
    nj = args.batch_jobs
    u = synthetic.uniform01(3, 4 * nj).reshape(nj, 4)
    pmax, smax = args.batch_pmax, args.batch_smax
    g2, gs2 = synthetic.double_gyre(512, 256, 51)
    g3, gs3 = synthetic.abc_flow(128, 26)
    specs = []
    for i in range(nj):
        dim = 2 if i % 2 == 0 else 3
        integ = 0 if u[i, 0] < 0.25 else 1
        n = int(round(math.exp(math.log(1e3) + u[i, 1] * (math.log(pmax) - math.log(1e3)))))
        steps = int(100 + u[i, 2] * (smax - 100))
        specs.append((dim, integ, n, steps))
    print(f"# batch: {nj} jobs, {sum(x[2] for x in specs)} particles, "
          f"{sum(x[2] * x[3] for x in specs):.3e} nominal particle-steps", file=sys.stderr, flush=True)
    seeds = []
    for i, (dim, integ, n, steps) in enumerate(specs):
        if dim == 2:
            seeds.append(synthetic.uniform_seeds(1000 + i, n, [0.01, 0.01], [1.99, 0.99]))
        else:
            seeds.append(synthetic.uniform_seeds(1000 + i, n, [0.0] * 3, [2 * math.pi] * 3))

    def desc(dim, integ):
        d = capi.JobDesc()
        d.dim, d.integrator, d.windfield_mode, d.dtd = dim, integ, 1, 1.0
        if dim == 2:
            d.dt = 0.01
            d.grid_scale[0], d.grid_scale[1], d.grid_scale[2] = gs2[0], gs2[1], gs2[0]
            d.domain[0] = float(np.float32(512) * np.float32(gs2[0]))
            d.domain[1] = float(np.float32(256) * np.float32(gs2[1]))
        else:
            d.dt = float(DT)
            for k in range(3):
                d.grid_scale[k] = gs3[k]
                d.domain[k] = float(np.float32(128) * np.float32(gs3[k]))
        return d

    out = {"workload": f"{nj} jobs (particles log-uniform [1e3, {pmax:.0e}], steps uniform [100, "
                       f"{smax}], half 2D double gyre / half 3D ABC, Euler:RK4 1:3)",
           "nominal_particle_steps": int(sum(x[2] * x[3] for x in specs))}
    # the sequential plain launches run first, with the context's default launch options (the
    # autotuner's choice for the batched form is not theirs)
    for form in ("sequential_plain_fp32", "batched_fp16"):
        dtype = capi.F16 if form == "batched_fp16" else capi.F32
        # plain launches sample slots separately and never reuse stencils: they run the batched
        # kernel's build without the stencil registers (more wavefronts per SIMD)
        ctx.set_option(capi.CTX_BATCH_REUSE, 1 if form == "batched_fp16" else 0)
        f2, f3 = make_field2(capi, ctx, g2, dtype), make_field(capi, ctx, g3, dtype)
        if form == "batched_fp16" and not args.no_autotune:
            tune = ctx.autotune(f3, desc(3, 1), n=65536, njobs=16, steps=50, trials=1)
            out["autotune"] = {"block": ctx.get_option(capi.CTX_BATCH_BLOCK),
                               "waves": ctx.get_option(capi.CTX_BATCH_WAVES),
                               "group": ctx.get_option(capi.CTX_BATCH_GROUP),
                               "best_probe_ms": min(t["ms"] for t in tune)}
        jobs = []
        for (dim, integ, n, steps), sd in zip(specs, seeds):
            j = capi.Job(ctx, desc(dim, integ), sd)
            if form != "batched_fp16":
                j.set_option(capi.OPT_PAIRED, 0)
            jobs.append((j, f2 if dim == 2 else f3, steps))
        if form == "batched_fp16":
            ctx.reorder([j for j, _, _ in jobs if j.n >= 4096])
        ctx.synchronize()
        ctx.take_work()
        ctx.take_kernel_ms()
        t0 = time.perf_counter()
        if form == "batched_fp16":
            ctx.batch_run([(j, f, dict(max_steps=st)) for j, f, st in jobs], timed=False)
        else:
            for j, f, st in jobs:
                ctx.batch_run([(j, f, dict(max_steps=st))], timed=False)
        ctx.synchronize()
        wall = time.perf_counter() - t0
        work = ctx.take_work()
        kms, nl = ctx.take_kernel_ms()
        out[form] = {"wall_s": wall, "device_s": kms / 1e3, "launches": nl,
                     "integrated_particle_steps": work, "integrated_particle_steps_per_s": work / wall,
                     "launch_options": {"block": ctx.get_option(capi.CTX_BATCH_BLOCK),
                                        "group": ctx.get_option(capi.CTX_BATCH_GROUP),
                                        "cu_waves": ctx.get_option(capi.CTX_BATCH_CU_WAVES),
                                        "reuse": ctx.get_option(capi.CTX_BATCH_REUSE)}}
        print(f"#   {form}: {wall:.2f} s wall, {kms / 1e3:.2f} s device, {work / wall:.3e} "
              f"integrated particle-steps/s", file=sys.stderr, flush=True)
        for j, _, _ in jobs:
            j.close()
        f2.close()
        f3.close()
    out["speedup"] = out["sequential_plain_fp32"]["wall_s"] / out["batched_fp16"]["wall_s"]
    print(json.dumps(out), flush=True)


def serving_child(duration):
    """The serving configuration (lambda = 50 jobs/s on four resident 256x256x96 FP16 datasets)
    for `duration` seconds of arrivals, run as `bench.py --workload serving` in a child process
    (its datasets are made with torch, whose HIP runtime must load before the engine's); returns
    its JSON record or the error."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--workload", "serving", "--serve-rate", "50",
           "--serve-duration", f"{duration:g}"]
    try:
        # the child's progress lines go straight to this process's stderr; its record to a pipe
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, text=True,
                           timeout=max(120.0, 6 * duration + 90))
    except subprocess.TimeoutExpired:
        return {"error": "timed out", "command": " ".join(cmd[1:])}
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"error": f"exit {r.returncode}", "command": " ".join(cmd[1:])}
    rec = json.loads(lines[-1])
    rec["command"] = "python3 bench.py " + " ".join(cmd[2:])
    return rec


def serving_workload(capi, ctx, args):
    """The serving configuration: Poisson arrivals (lambda jobs/s over a duration) onto four
    resident atmosphere-like datasets (256x256x96, stretched z, 48 hourly slices, FP16, seeds
    1-4); sizes as in the batched configuration (particles log-uniform [1e3, 4e6], steps
    uniform [100, 5000]), RK4:Euler 3:1, dt 60 s; seeds uniform in x-y and log-uniform in height,
    made on the device at admission. Jobs are admitted by HBM budget and advanced in rounds of
    the persistent batched kernel. Prints one JSON line with p50/p99 latency and throughput; with
    --serve-sweep, one run per listed rate on the same resident datasets and the saturation rate
    (the highest listed rate whose queue drained within 5 % of the arrival window)."""
    import torch
    from pygranite_amd import serving, synthetic
    dev = torch.device("cuda")
    nx = ny = 256
    nz, S = 96, 48
    ds = {}
    for k in range(4):
        wind_t, gs, zlev = synthetic.atmosphere(nx, ny, nz, S, seed=k + 1, device=dev)
        f = capi.Field(ctx, nx, ny, nz, 3, dtype=capi.F16, nslots=S, address=0)
        f.set_z_levels(zlev)
        torch.cuda.synchronize()  # the engine's copies are not ordered after torch's stream
        for sl in range(S):
            f.upload_device(sl, [wind_t[sl, c].data_ptr() for c in range(3)])
        ctx.synchronize()
        del wind_t
        dom = [float(np.float32(nx) * np.float32(gs[0])), float(np.float32(ny) * np.float32(gs[1])),
               float(zlev[-1])]
        ds[f"d{k}"] = serving.Dataset(f, gs, dom, 3600.0, S, zlev)
    torch.cuda.empty_cache()
    workload = ("4 datasets 256x256x96 FP16 x 48 slices; Poisson arrivals; particles "
                f"log-uniform [1e3, {args.batch_pmax:.0e}], steps uniform [100, {args.batch_smax}]")

    def one_rate(rate, duration):
        arr = serving.poisson_arrivals(4, rate, duration)
        u = synthetic.uniform01(5, 5 * len(arr)).reshape(len(arr), 5)
        reqs = []
        for i, t in enumerate(arr):
            n = int(round(math.exp(math.log(1e3) + u[i, 0] * (math.log(args.batch_pmax) - math.log(1e3)))))
            dsn = f"d{int(u[i, 1] * 4)}"
            dom, z1 = ds[dsn].domain, float(ds[dsn].zlevels[1])

            def seed_fn(n=n, dom=dom, z1=z1, s=int(u[i, 4] * 2**31)):
                g = torch.Generator(device=dev)
                g.manual_seed(s)
                r = torch.rand((n, 3), generator=g, device=dev, dtype=torch.float64)
                xyz = torch.empty((n, 3), device=dev, dtype=torch.float32)
                xyz[:, 0] = (r[:, 0] * dom[0]).float()
                xyz[:, 1] = (r[:, 1] * dom[1]).float()
                xyz[:, 2] = (z1 * torch.exp(r[:, 2] * math.log(dom[2] / z1))).float()
                return xyz

            reqs.append(serving.Request(rid=i, arrival=float(t), dataset=dsn, seeds=None, n=n,
                                        seed_fn=seed_fn, steps=int(100 + u[i, 2] * (args.batch_smax - 100)),
                                        integrator=0 if u[i, 3] < 0.25 else 1, dt=60.0))
        offered = sum(r.n * r.steps for r in reqs) / duration
        print(f"# serving: lambda {rate:g}/s, {len(reqs)} requests over {duration} s, offered "
              f"{offered:.3e} nominal particle-steps/s", file=sys.stderr, flush=True)
        srv = serving.Server(ctx, ds, hbm_budget=int(args.serve_budget_gb * (1 << 30)), quantum=args.serve_quantum,
                             reorder_every=args.serve_reorder_every)
        ctx.take_work()
        srv.run(reqs, progress=lambda m: print("#   " + m, file=sys.stderr, flush=True))
        rep = srv.report()
        rep["integrated_particle_steps"] = ctx.take_work()
        rep["integrated_particle_steps_per_s"] = rep["integrated_particle_steps"] / rep["wall_s"]
        rep.update(arrival_rate_per_s=rate, duration_s=duration,
                   regroup_every_rounds=args.serve_reorder_every,
                   offered_nominal_particle_steps_per_s=offered,
                   drain_after_last_arrival_s=rep["wall_s"] - float(arr[-1]) if len(arr) else 0.0,
                   kept_up=bool(rep["wall_s"] <= 1.05 * duration))
        return rep

    if args.serve_sweep:
        rates = [float(x) for x in args.serve_sweep.split(",")]
        reps = [one_rate(r, args.serve_duration) for r in rates]
        ok = [r["arrival_rate_per_s"] for r in reps if r["kept_up"]]
        print(json.dumps({"workload": workload, "runs": reps, "saturation_rate_per_s": max(ok) if ok else None,
                          "saturation_rule": "highest swept rate whose run ended within 5 % of its arrival window"}),
              flush=True)
    else:
        rep = one_rate(args.serve_rate, args.serve_duration)
        rep["workload"] = workload
        print(json.dumps(rep), flush=True)
    for d in ds.values():
        d.field.close()


def e2e_workload(args):
    """End-to-end through the drop-in API (TrajectoryIntegrator(...).compute()), against the
    kernel-only device time of the same call (TrajectorySet.deviceTime()):
      * the 3D ABC configuration with every step saved (positions, curvature, expression) on a
        stated particle subsample: upload, integration, saved columns streamed to the host while
        integration continues, and the (N, T, D) export;
      * the atmosphere configuration (512x512x64 stretched z, 25 hourly slices handed over one
        per window by a Python DataLoader, FP16 storage, Kahan + FP64 time, expression):
        per-window uploads overlapping integration, final positions only.
    Prints one JSON line."""
    import torch
    torch.zeros(1, device="cuda")
    import pygranite_amd as pg
    from pygranite_amd import synthetic
    Loader = synthetic.make_loader_class(pg.DataLoader)
    out = {}

    def integrated(res, steps):
        st = np.array([int(r.value) for r in res.abortReasons()])
        ab = res.abortSteps().astype(np.int64)
        return int(np.where(st == 0, steps, ab + 1).sum())

    wind, gs = synthetic.abc_flow(128, 16)
    n = args.e2e_particles
    seeds = np.ascontiguousarray(synthetic.uniform_seeds(1, 1 << 20, [0.0] * 3, [2 * math.pi] * 3)[:n])
    cases = [("abc_float16", "Float16", None), ("abc_float32", "Float32", None),
             ("abc_float32_regroup_edges_only", "Float32", 0)]
    for name, storage, regroup in cases:
        st = pg.IntegratorSettings()
        st.Space, st.Integrator = pg.Space.Space3D, pg.Integrator.ClassicRungeKutta
        st.DeltaT, st.DataTimeDistance, st.GridScale, st.Offset = 0.005, 1.0, gs, [0.0, 0.0, 0.0]
        st.MaximumSimulationTime, st.SaveInterval = 10.0, 1
        st.CurvatureMode = pg.CurvatureMode.IndividualAndTotalCurvature
        st.AdditionalCompute = ["sqrt(u*u+v*v+w*w)"]
        st.WindfieldStorage = getattr(pg.StorageType, storage)
        if regroup is not None:
            st.RegroupInterval = regroup  # regroup at window edges only
        t0 = time.perf_counter()
        integ = pg.TrajectoryIntegrator(st, Loader(wind), pg.TrajectorySet(seeds))
        t1 = time.perf_counter()
        res = integ.compute()
        t2 = time.perf_counter()
        traj = res.trajectories()
        curv = res.curvatures()
        ex = res.computeInfo(0)
        t3 = time.perf_counter()
        nbytes = traj.nbytes + curv.nbytes + ex.nbytes
        work = integrated(res, 2000)
        out[name] = {
            "particles": n, "steps": 2000, "saved_columns": int(traj.shape[1]),
            "regroup_interval": int(st.RegroupInterval),
            "output_bytes": int(nbytes), "construct_s": t1 - t0, "compute_s": t2 - t1,
            "device_s": res.deviceTime(), "export_s": t3 - t2,
            "integrated_particle_steps": work,
            "end_to_end_particle_steps_per_s": work / (t3 - t0),
            "compute_particle_steps_per_s": work / (t2 - t1),
            "kernel_only_particle_steps_per_s": work / max(1e-9, res.deviceTime()),
            "output_GB_per_s_end_to_end": nbytes / (t3 - t0) / 1e9}
        print(f"#   e2e {name}: construct {t1 - t0:.2f} s, compute {t2 - t1:.2f} s (device "
              f"{res.deviceTime():.3f} s), export {t3 - t2:.2f} s, {nbytes / 1e9:.1f} GB out",
              file=sys.stderr, flush=True)
        del traj, curv, ex, res, integ
    if not args.e2e_no_atmosphere:
        nx = ny = args.atmo_nx
        nz, S = 64, 25
        t0 = time.perf_counter()
        wind_t, gs_a, zlev = synthetic.atmosphere(nx, ny, nz, S, seed=1, device=torch.device("cuda"))
        host = [wind_t[s].cpu().numpy() for s in range(S)]
        del wind_t
        torch.cuda.empty_cache()
        slices = np.stack(host)
        del host
        print(f"# e2e atmosphere field on the host in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        dom = [float(np.float32(nx) * np.float32(gs_a[0])), float(np.float32(ny) * np.float32(gs_a[1]))]
        na = args.atmo_particles
        xy = synthetic.uniform_seeds(21, na, [0.0, 0.0], dom)
        z = synthetic.loguniform_heights(22, na, float(zlev[1]), float(zlev[-1]))
        seeds_a = np.ascontiguousarray(np.concatenate([xy, z[:, None]], axis=1), np.float32)
        st = pg.IntegratorSettings()
        st.Space, st.Integrator = pg.Space.Space3D, pg.Integrator.ClassicRungeKutta
        st.DeltaT, st.DataTimeDistance, st.GridScale, st.Offset = 60.0, 3600.0, list(gs_a), [0.0, 0.0, 0.0]
        st.MaximumSimulationTime, st.SaveInterval = 1440 * 60.0, 0
        st.KahanSummation, st.TimeFP64, st.AllowAbs = True, True, True
        st.WindfieldStorage = pg.StorageType.Float16
        st.ZLevels = [float(v) for v in zlev]
        st.AdditionalCompute = ["w*1000/(abs(u)+abs(v)+1)"]
        t0 = time.perf_counter()
        integ = pg.TrajectoryIntegrator(st, Loader(slices), pg.TrajectorySet(seeds_a))
        t1 = time.perf_counter()
        res = integ.compute()
        t2 = time.perf_counter()
        pos = res.trajectories()
        t3 = time.perf_counter()
        work = integrated(res, 1440)
        out["atmosphere_fp16"] = {
            "particles": na, "steps": 1440, "windows": S - 1, "slice_bytes_fp32": int(slices[0].nbytes),
            "construct_s": t1 - t0, "compute_s": t2 - t1, "device_s": res.deviceTime(), "export_s": t3 - t2,
            "integrated_particle_steps": work,
            "end_to_end_particle_steps_per_s": work / (t3 - t0),
            "compute_particle_steps_per_s": work / (t2 - t1),
            "kernel_only_particle_steps_per_s": work / max(1e-9, res.deviceTime())}
        print(f"#   e2e atmosphere: construct {t1 - t0:.2f} s, compute {t2 - t1:.2f} s (device "
              f"{res.deviceTime():.3f} s), export {t3 - t2:.2f} s", file=sys.stderr, flush=True)
        del pos
    print(json.dumps(out), flush=True)


def copy_calibration():
    """Three copies of a 1 GiB device buffer: the reference for scaling HBM byte counters."""
    import torch
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copies": 3, "bytes_read_each": a.numel() * 4, "bytes_written_each": a.numel() * 4}))


def load_counters(form):
    """Per-launch hardware counters of a form's step kernel for the default command, from the
    committed rocprofv3 --pmc summary (profiles/counters_<form>_abc3d.json): HBM bytes
    (FETCH_SIZE and WRITE_SIZE scaled by the copy calibration) and VALU wave instructions. None
    when absent."""
    p = os.path.join(ROOT, "profiles", f"counters_{form}_abc3d.json")
    if not os.path.exists(p):
        return None
    with open(p) as fh:
        c = json.load(fh)
    c["source"] = os.path.relpath(p, ROOT) + " (" + c.get("command", "rocprofv3 --pmc") + ")"
    return c


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40,
                    help="timed launches; each runs ceil(total / steps) integration steps, so the "
                         "timed region is always the whole job")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--chunk", type=int, default=0,
                    help="integration steps per launch (0: ceil(total-steps / steps))")
    ap.add_argument("--total-steps", type=int, default=2000, help="steps of the headline job")
    ap.add_argument("--repeats", type=int, default=5,
                    help="timed runs of the whole job per form (value: the median)")
    ap.add_argument("--error-particles", type=int, default=16384,
                    help="particles compared with the FP64 oracle")
    ap.add_argument("--particles", type=int, default=1 << 20)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--jobs", type=int, default=16)
    ap.add_argument("--reorder-every", type=int, default=1,
                    help="regroup particles by cell every K launches as well as at window edges "
                         "(0: at window edges only)")
    ap.add_argument("--edge-regroup-min", type=int, default=4096,
                    help="regroup at window edges only while at least this many particles are live")
    ap.add_argument("--regroup-min", type=int, default=65536,
                    help="regroup every K launches only while the live range (all jobs) holds at "
                         "least this many particles; smaller launches are latency-bound")
    ap.add_argument("--quad", type=int, default=-1,
                    help="single-job forms: live-particle limit for quad launches (GR_OPT_QUAD; -1: the "
                         "library default)")
    ap.add_argument("--forms", default="plain_fp32,fp32_local,fp32k,fp16,fp16_batched,fp32k_batched,fp32_tile,fp16_tile")
    ap.add_argument("--tile-bytes", type=int, default=32768,
                    help="LDS field tile per workgroup of the *_tile forms (GR_OPT_TILE)")
    ap.add_argument("--tile-halo", type=int, default=1, help="cells around the tile's box (GR_OPT_TILE_HALO)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the lite batch/serving runs")
    ap.add_argument("--serve-default-duration", type=float, default=20.0,
                    help="arrival window of the serving configuration run in the default command "
                         "(lambda 50 jobs/s on the four atmosphere datasets; 0: skip)")
    ap.add_argument("--no-gyre", action="store_true", help="skip the double-gyre GPU / CPU legs")
    ap.add_argument("--workload", default="abc",
                    choices=["abc", "atmosphere", "copy", "batch", "serving", "e2e"],
                    help="copy: a plain 1 GiB device copy (x3), the calibration run for "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--atmo-nx", type=int, default=512)
    ap.add_argument("--e2e-particles", type=int, default=1 << 18,
                    help="e2e workload: particles of the ABC run with every step saved")
    ap.add_argument("--e2e-no-atmosphere", action="store_true")
    ap.add_argument("--atmo-particles", type=int, default=1 << 24)
    ap.add_argument("--atmo-error-particles", type=int, default=4096,
                    help="particles of the atmosphere workload checked against the FP64 oracle")
    ap.add_argument("--batch-block", type=int, default=0, help="batched kernel workgroup size (0: default)")
    ap.add_argument("--batch-waves", type=int, default=-1, help="batched kernel register budget (0 or 4)")
    ap.add_argument("--batch-group", type=int, default=0,
                    help="batched kernel compute units per queue (0: default / autotuned)")
    ap.add_argument("--batch-cu-waves", type=int, default=-1,
                    help="batched kernel wavefronts per CU (0: register-limited; default: autotuned)")
    ap.add_argument("--batch-uniform", type=int, default=-1,
                    help="per-variant batched kernels for stencil-reuse launches of one variant (0 or 1)")
    ap.add_argument("--reuse", type=int, default=-1,
                    help="stencil-reuse build for every launch: batched (context) and single-job forms")
    ap.add_argument("--cu-waves", type=int, default=0,
                    help="single-job forms: wavefronts per CU of the step kernel (0: register-limited)")
    ap.add_argument("--batch-jobs", type=int, default=256)
    ap.add_argument("--batch-pmax", type=float, default=4e6)
    ap.add_argument("--batch-smax", type=int, default=5000)
    ap.add_argument("--serve-rate", type=float, default=50.0)
    ap.add_argument("--serve-sweep", default="",
                    help="comma-separated arrival rates: one serving run each and the saturation rate")
    ap.add_argument("--serve-duration", type=float, default=120.0)
    ap.add_argument("--serve-quantum", type=int, default=250)
    ap.add_argument("--serve-budget-gb", type=float, default=64.0)
    ap.add_argument("--serve-reorder-every", type=int, default=1,
                    help="serving rounds between joint regroupings of the running jobs")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the batched-kernel autotuner (default: tune before the batched form)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the path would shard by "
              "particles (each GPU holding the window's slices and N/G particles, no exchange "
              "during integration), and no run on several GPUs is measured here", file=sys.stderr)
        sys.exit(2)

    if args.workload == "copy":
        copy_calibration()
        return
    if args.workload == "e2e":
        e2e_workload(args)
        return
    if args.workload in ("atmosphere", "serving"):
        # torch generates fields and seeds on the device and hands their pointers to the engine:
        # its HIP runtime must be the process's runtime, so it initialises before the engine
        import torch
        torch.zeros(1, device="cuda")
    from pygranite_amd import capi, synthetic
    if capi.device_count() == 0:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    ctx = capi.Context(0)
    info = ctx.info()
    if args.batch_block:
        ctx.set_option(capi.CTX_BATCH_BLOCK, args.batch_block)
    if args.batch_waves >= 0:
        ctx.set_option(capi.CTX_BATCH_WAVES, args.batch_waves)
    if args.batch_cu_waves >= 0:
        ctx.set_option(capi.CTX_BATCH_CU_WAVES, args.batch_cu_waves)
    if args.batch_group:
        ctx.set_option(capi.CTX_BATCH_GROUP, args.batch_group)
    if args.reuse >= 0:
        ctx.set_option(capi.CTX_BATCH_REUSE, args.reuse)
    if args.batch_uniform >= 0:
        ctx.set_option(capi.CTX_BATCH_UNIFORM, args.batch_uniform)
    if args.workload == "atmosphere":
        atmosphere_workload(capi, ctx, args)
        return
    if args.workload == "batch":
        batch_workload(capi, ctx, args)
        return
    if args.workload == "serving":
        serving_workload(capi, ctx, args)
        return
    n = args.grid
    total = args.total_steps
    chunk = args.chunk or -(-total // max(1, args.steps))
    wind, gs = synthetic.abc_flow(n, 16)
    seeds = synthetic.uniform_seeds(1, args.particles, [0.0] * 3, [2 * math.pi] * 3)
    print(f"# device {info['name']} CUs={info['compute_units']}", file=sys.stderr, flush=True)
    f32 = make_field(capi, ctx, wind, capi.F32)
    f16 = make_field(capi, ctx, wind, capi.F16)

    forms = args.forms.split(",")
    tune = {}
    fixed = bool(args.no_autotune or args.batch_block or args.batch_waves >= 0 or args.batch_cu_waves >= 0
                 or args.batch_group)

    def autotune(name):
        """Launch configuration of the batched kernel for one form, from the autotuner: probe jobs
        of the same count, size and description over the form's own field, one launch, before
        any timing of the form."""
        kahan = name == "fp32k_batched"
        t = ctx.autotune(f32 if kahan else f16, build_desc(capi, gs, n, kahan=kahan),
                         n=max(1, args.particles // args.jobs), njobs=args.jobs, steps=min(chunk, 100), trials=2)
        chosen = {"block": ctx.get_option(capi.CTX_BATCH_BLOCK), "group": ctx.get_option(capi.CTX_BATCH_GROUP),
                  "cu_waves": ctx.get_option(capi.CTX_BATCH_CU_WAVES), "waves": ctx.get_option(capi.CTX_BATCH_WAVES)}
        print(f"# autotune {name}: " + ", ".join(f"b{c['block']}/g{c['group']}/cw{c['cu_waves']} {c['ms']:.3f} ms"
                                                for c in t)
              + f" -> block {chosen['block']}, group {chosen['group']}, CU waves {chosen['cu_waves']}",
              file=sys.stderr, flush=True)
        tune[name] = {"candidates": t, "chosen": chosen}

    def autotune_single(name, f, kahan):
        """Launch configuration of a single-job form from the single-job autotuner (workgroup
        size x lanes per particle x LDS field tile) on a probe job of the form's description
        and particle count over the form's own field, before any timing of the form."""
        probe = capi.Job(ctx, build_desc(capi, gs, n, kahan=kahan), seeds, save_capacity=1)
        t = probe.autotune(f, steps=min(chunk, 100), trials=2)
        probe.close()
        best = min(t, key=lambda c: c["ms"])
        print(f"# autotune {name}: " + ", ".join(f"b{c['block']}/l{c['lanes']}/t{c['tile_bytes'] // 1024}k "
                                                f"{c['ms']:.3f} ms" for c in t)
              + f" -> block {best['block']}, lanes {best['lanes']}, tile {best['tile_bytes']}",
              file=sys.stderr, flush=True)
        tune[name] = {"candidates": t, "chosen": best}
        return dict(block=best["block"], tile=best["tile_bytes"], quad=(1 << 29) - 1 if best["lanes"] == 4 else -1)

    res = {}
    for name in forms:
        reps = args.repeats
        print(f"# form {name}: {args.warmup}+{args.steps} launches of {chunk} steps ({total}-step job, "
              f"{reps} timed run(s)), {args.particles} particles", file=sys.stderr, flush=True)
        kw = dict(total=total, repeats=reps)
        if name == "plain_fp32":
            r = run_single(capi, ctx, f32, seeds, gs, n, chunk, args.steps, args.warmup, False, False, **kw)
        elif name == "fp32_local":
            kw["cu_waves"] = args.cu_waves
            kw["quad"] = args.quad
            if not args.no_autotune:
                kw.update(autotune_single(name, f32, False))
            r = run_single(capi, ctx, f32, seeds, gs, n, chunk, args.steps, args.warmup, True, True,
                           args.reorder_every, args.regroup_min, edge_min=args.edge_regroup_min, **kw)
        elif name in ("fp32_tile", "fp16_tile"):
            kw["cu_waves"] = args.cu_waves
            kw["quad"] = args.quad
            r = run_single(capi, ctx, f32 if name == "fp32_tile" else f16, seeds, gs, n, chunk, args.steps,
                           args.warmup, True, True, args.reorder_every, args.regroup_min,
                           tile=args.tile_bytes, tile_halo=args.tile_halo, edge_min=args.edge_regroup_min, **kw)
        elif name in ("fp16", "fp32k"):
            kw["cu_waves"] = args.cu_waves
            kw["reuse"] = args.reuse
            kw["quad"] = args.quad
            if not args.no_autotune:
                kw.update(autotune_single(name, f16 if name == "fp16" else f32, name == "fp32k"))
            r = run_single(capi, ctx, f16 if name == "fp16" else f32, seeds, gs, n, chunk, args.steps,
                           args.warmup, True, True, args.reorder_every, args.regroup_min,
                           kahan=name == "fp32k", edge_min=args.edge_regroup_min, **kw)
        elif name in ("fp16_batched", "fp32k_batched"):
            if not fixed:
                autotune(name)
            r = run_batched(capi, ctx, f16 if name == "fp16_batched" else f32, seeds, gs, n, chunk, args.steps,
                            args.warmup, args.jobs, args.reorder_every, args.regroup_min,
                            kahan=name == "fp32k_batched", edge_min=args.edge_regroup_min, **kw)
        else:
            raise SystemExit(f"unknown form {name}")
        r["options"] = {"block": ctx.get_option(capi.CTX_BATCH_BLOCK), "group": ctx.get_option(capi.CTX_BATCH_GROUP),
                        "cu_waves": ctx.get_option(capi.CTX_BATCH_CU_WAVES),
                        "waves": ctx.get_option(capi.CTX_BATCH_WAVES)}
        r["nominal_psps"] = args.particles * r["steps_done"] / r["wall_s"]
        r["psps"] = r["work"] / r["wall_s"]  # integrated particle-steps per second
        r["avg_launch_ms"] = r["kernel_ms"] / max(1, r["launches"])
        res[name] = r
        print(f"#   {name}: {r['psps']:.3e} integrated particle-steps/s (wall, median; "
              f"{r['work'] / (args.particles * max(1, r['steps_done'])):.3f} of nominal), device "
              f"{r['kernel_ms']:.1f} ms in {r['launches']} launches, {r['steps_done']} steps",
              file=sys.stderr, flush=True)

    err = None
    if not args.no_error:
        print(f"# FP64 oracle error check ({args.error_particles} particles)", file=sys.stderr, flush=True)
        sub = min(args.error_particles, args.particles)
        cmp_forms = {k: res[k] for k in FORMS if k in res}
        rounders = {k: synthetic.round_fp16 for k in ("fp16", "fp16_batched", "fp16_tile")}
        # every form is integrated over the same number of steps
        err = fp64_error(wind, seeds[:sub], gs, res[forms[0]]["steps_done"], cmp_forms, rounders,
                         kahan_forms=("fp32k", "fp32k_batched"))
    # the headline: the fastest form (plain_fp32, the reference's own form, is the baseline)
    # whose RMS position error against the FP64 oracle meets the 1e-4 (domain-relative) bound,
    # with every termination mismatch explained (diverged paths or a threshold decided within
    # one step); the fastest form when no error was measured
    cands = [k for k in res if k != "plain_fp32"] or list(res)
    ok = [k for k in cands if err and err.get(k, {}).get("rms") is not None and err[k]["rms"] <= ERR_TARGET
          and err[k].get("mismatch_other", 0) == 0]
    head = max(ok or cands, key=lambda k: res[k]["psps"])
    hv = res[head]
    outputs = 3 * 4 + 4 + 4  # saved position, curvature, expression per particle-step
    # + Kahan compensation read and written with the positions
    B16 = algorithmic_bytes(2 if "fp16" in head else 4, outputs) + (2 * 3 * 4 if head in KAHAN_FORMS else 0)
    # algorithmic bytes of the integrated particle-steps over the kernel's device time, both
    # from this run (HIP events around every step-kernel launch of the timed job)
    ps_per_launch = hv["work"] / max(1, hv["launches"])
    achieved = ps_per_launch * B16 / (hv["avg_launch_ms"] / 1e3) / 1e9

    extras = {}
    if not args.no_extras:
        print("# lite batched-jobs run", file=sys.stderr, flush=True)
        extras["batched_jobs"] = batched_jobs_lite(capi, ctx)
        print("# lite serving run", file=sys.stderr, flush=True)
        extras["serving"] = serving_lite(capi, ctx)
    if not args.no_extras and args.serve_default_duration > 0:
        print(f"# serving configuration: lambda 50 jobs/s for {args.serve_default_duration:g} s (child process)",
              file=sys.stderr, flush=True)
        extras["serving_configuration"] = serving_child(args.serve_default_duration)
    gyre = None
    if not args.no_gyre:
        print("# double-gyre configuration: CPU (all threads, one thread) and GPU", file=sys.stderr, flush=True)
        gcpu = gyre_cpu_baseline() if not args.no_cpu else None
        ggpu = gyre_gpu(capi, ctx, gcpu["saved"] if gcpu else None)
        gyre = {"workload": "2D double gyre 512x256 x 11 slices, 65,536 particles, Euler dt 0.01, "
                            "1000 steps, saved every 10th step", "gpu": ggpu}
        if gcpu:
            gcpu.pop("saved")
            gyre["cpu"] = gcpu
            gyre["gpu_vs_cpu_all_threads"] = (ggpu["fp32_local"]["particle_steps_per_s"]
                                              / gcpu["all"]["particle_steps_per_s"])
    cpu = None
    if not args.no_cpu:
        print("# CPU baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(wind, gs)
        cpu["cpu_model"] = cpu_model()
    counters = load_counters(head)
    # the measured counterpart of the published peak: a 1 GiB device-to-device streaming copy
    copy_gbs = ctx.copy_bandwidth(1 << 30, 5)

    plain = res.get("plain_fp32", {}).get("psps")
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": counters.get("hbm_bytes_per_launch") if counters else None,
            "achieved_is": "algorithmic bytes (stencil gathers at the storage width, state and "
                           "saved outputs) of the launch's integrated particle-steps over its "
                           "average device time; most gathers are served by L2 / MALL",
            "kernel": ("k_batch (RK4 3D " + ("FP16" if "fp16" in head else "FP32 + Kahan") + " items)")
                      if head.endswith("_batched") else
                      "k_advance (RK4 3D " + ("FP16" if "fp16" in head else "FP32" + (" + Kahan" if head in KAHAN_FORMS else ""))
                      + ", time-paired; stencil-reuse and quad builds)",
            "peak_measured_copy_GBs": copy_gbs,
            "bytes_per_particle_step": B16, "avg_launch_ms": hv["avg_launch_ms"],
            "launches": hv["launches"], "integrated_particle_steps_per_launch": ps_per_launch,
            "algorithmic_bytes_per_launch": ps_per_launch * B16}
    if counters:
        # measured per-launch counters of the same launches (rocprofv3 --pmc restricted to the
        # step kernel, summarised under profiles/): HBM bytes and VALU instructions, scaled by the
        # integrated particle-steps per launch when this run's launches did other work
        scale = ps_per_launch / counters["integrated_particle_steps_per_launch_pmc_run"]
        traffic = counters["hbm_bytes_per_launch"] * scale
        roof["traffic"] = traffic
        roof["traffic_source"] = counters.get("source")
        roof["traffic_scaled_by"] = scale
        roof["hbm_actual_GBs"] = traffic / (hv["avg_launch_ms"] / 1e3) / 1e9
        roof["hbm_actual_frac"] = roof["hbm_actual_GBs"] / HBM_PEAK_GBS
        roof["traffic_over_algorithmic"] = traffic / (ps_per_launch * B16)
        valu = counters["valu_insts_per_launch"] * scale
        rate = valu / (hv["avg_launch_ms"] / 1e3)
        dense = counters["per_launch"][0]
        roof["valu"] = {"wave_insts_per_launch": valu, "achieved_G_per_s": rate / 1e9,
                        "peak_G_per_s": VALU_PEAK_WAVE_INSTS / 1e9, "frac": rate / VALU_PEAK_WAVE_INSTS,
                        "frac_first_launch": dense["valu_frac"],
                        "wave_insts_per_particle_step": counters["valu_insts_per_particle_step"],
                        "peak_is": "256 CUs x 2.4 GHz x one wave64 VALU instruction per clock"}
        roof["binding"] = ("VALU issue in the dense launches (frac_first_launch), one wavefront's "
                           "dependent chain in the tail launches; HBM carries traffic / avg launch "
                           "time = hbm_actual_GBs")
    line = {
        "metric": "RK4 3D particle-steps/s vs plain FP32 at <=1e-4 RMS err; p50/p99 job latency",
        "value": hv["psps"],
        "nominal_particle_steps_per_s": hv["nominal_psps"],
        "integrated_fraction": hv["work"] / (args.particles * max(1, hv["steps_done"])),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": hv["wall_s"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (analytic ABC flow, SplitMix64 seeds); value counts integrated "
                "particle-steps (steps begun alive), counted on the device; median of "
                f"{len(hv['repeat_walls_s'])} timed runs of the whole job",
        "config": {"workload": f"3D ABC flow RK4, {n}^3 x 16 slices, {args.particles:,} particles, dt "
                               f"0.005, {hv['steps_done']} steps, curvature + sqrt(u*u+v*v+w*w) saved "
                               "every step",
                   "form": head, "storage": "fp16" if head.startswith("fp16") else "fp32",
                   "kahan": head in KAHAN_FORMS, "error_target_rms": ERR_TARGET,
                   # RMS over the particles whose termination agrees with the FP64 oracle, and no
                   # mismatch other than diverged paths or a threshold decided within one step
                   "meets_error_target": bool(err and err.get(head, {}).get("rms") is not None
                                              and err[head]["rms"] <= ERR_TARGET
                                              and err[head].get("mismatch_other", 0) == 0),
                   "particles": args.particles, "grid": n, "steps_done": hv["steps_done"],
                   "steps_per_launch": chunk, "jobs": args.jobs if head.endswith("_batched") else 1,
                   "regroup_every_launches": args.reorder_every, "regroup_min_live": args.regroup_min,
                   "edge_regroup_min_live": args.edge_regroup_min,
                   "batch_block": hv["options"]["block"], "batch_waves": hv["options"]["waves"],
                   "batch_group": hv["options"]["group"], "batch_cu_waves": hv["options"]["cu_waves"],
                   "single_job_cu_waves": args.cu_waves},
        "forms_psps": {k: v["psps"] for k, v in res.items()},
        "forms_device_ms": {k: v["kernel_ms"] for k, v in res.items()},
        "tile_workgroup_steps": {k: v["tile_workgroup_steps"] for k, v in res.items() if "tile_workgroup_steps" in v},
        "forms_steps_done": {k: v["steps_done"] for k, v in res.items()},
        "repeat_walls_s": hv["repeat_walls_s"],
        "autotune": tune,
        "plain_fp32_value": plain,
        "speedup_vs_plain_fp32": hv["psps"] / plain if plain else None,
        "fp64_error_domain_relative": err,
        "roofline": roof,
        "cpu_baseline": cpu,
        "double_gyre": gyre,
        "extras": extras,
        "device": info["name"],
    }
    print(json.dumps(line), flush=True)

if __name__ == "__main__":
    main()
