#!/usr/bin/env python3
"""Benchmark: RK4 3D particle-steps/s on the time-modulated ABC flow (one MI355X).

Workload (the 3D ABC configuration): 128^3 grid over [0, 2*pi)^3, 16 hourly-spaced slices
(DataTimeDistance 1), 1,048,576 uniform seeds (SplitMix64 seed 1), RK4 with dt = 0.005,
per-step curvature (IndividualAndTotalCurvature) and the expression sqrt(u*u+v*v+w*w), saved
at every step into a device-resident column ring. All 16 slices are resident in HBM before the
timed region; a bench "step" is one launch of the fused step kernel advancing every particle by
`--chunk` integration steps (default 50, so 40 bench steps = the full 2000-step job).

Two forms run on identical inputs: the plain FP32 form (FP32 slices) and the FP16 form (FP16
slices, FP32 interpolation and positions). `value` is the FP16 form; the FP32 figure, the
speed-up and the FP16 form's RMS / max position error against the FP64 CPU oracle (as a fraction
of the domain extent 2*pi, on a particle subsample) are reported beside it.
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


def build_desc(capi, dim, gs, dt, n_prog_src, curvature=3, save_interval=1):
    d = capi.JobDesc()
    d.dim = dim
    d.integrator = 1
    d.curvature = curvature
    d.windfield_mode = 1
    d.dt = dt
    d.dtd = 1.0
    for k in range(3):
        d.grid_scale[k] = gs[k]
        d.domain[k] = 128 * gs[k]
        d.offset[k] = 0.0
    d.save_interval = save_interval
    progs = [capi.compile_expr(s, dim=dim, curvature=curvature) for s in n_prog_src]
    d.nprog = len(progs)
    for i, p in enumerate(progs):
        d.prog[i] = p
    return d


def algorithmic_bytes(storage_bytes: int, outputs: int) -> int:
    """Bytes per RK4-3D particle-step: 4 stages x 2 slices x 3 comps x 8 corners x s
    + position read/write and status (2*3*4 + 4) + saved outputs."""
    return 4 * 2 * 3 * 8 * storage_bytes + (2 * 3 * 4 + 4) + outputs


def run_form(capi, ctx, wind, seeds, gs, dtype, chunk, steps, warmup, exprs, paired=True,
             reorder=True):
    S = wind.shape[0]
    n = wind.shape[2]
    f = capi.Field(ctx, n, n, n, 3, dtype=dtype, nslots=S, address=0)
    for s in range(S):
        f.upload(s, [wind[s, c] for c in range(3)])
    for s in range(S):
        f.wait(s)
    dt = np.float32(0.005)
    desc = build_desc(capi, 3, gs, float(dt), exprs)
    job = capi.Job(ctx, desc, seeds, save_capacity=chunk)
    job.set_option(capi.OPT_PAIRED, 1 if paired else 0)
    if reorder:
        job.reorder()
    ctx.synchronize()

    counter = 0
    t32 = np.float32(0.0)
    window = 0
    max_wt = np.float32(1.0)
    kernel_ms = []

    def one_step():
        nonlocal counter, t32, window, max_wt
        # split the chunk at window edges exactly like the host schedule (FP32 time)
        left = chunk
        while left:
            k = 0
            t = t32
            done = False
            while k < left:
                k += 1
                t = np.float32(t + dt)
                if t >= max_wt:
                    done = True
                    break
            w = capi.Window()
            w.wind = f.h
            w.wind_back, w.wind_front = window, window + 1
            w.time_back, w.time_front = float(np.float32(window)), float(np.float32(window + 1))
            w.time_back64, w.time_front64 = float(window), float(window + 1)
            w.time_start = float(t32)
            w.time_start64 = float(t32)
            w.counter_start = counter
            w.nsteps = k
            job.advance(w)
            kernel_ms.append(job.last_advance_ms())
            counter += k
            t32 = t
            left -= k
            if done:
                if reorder:
                    job.reorder()  # re-sort by cell at every window edge (inside the timing)
                window += 1
                max_wt = np.float32(max_wt + np.float32(1.0))
                if window + 1 >= S:
                    raise RuntimeError("ran out of slices")

    for _ in range(warmup):
        one_step()
    kernel_ms.clear()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        one_step()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    pos = job.read(capi.OUT_POSITIONS)
    status = job.read(capi.OUT_STATUS)
    result = {
        "wall_s": wall,
        "kernel_ms": float(np.sum(kernel_ms)),
        "launches": len(kernel_ms),
        "steps_done": counter,
        "positions": pos,
        "status": status,
        "field_bytes": f.bytes_per_slot() * S,
    }
    job.close()
    f.close()
    return result


def fp64_error(wind, seeds_sub, gs, steps_done, got_sub):
    import oracle
    st = dict(Space=0, Integrator=1, DeltaT=0.005, DataTimeDistance=1.0, GridScale=gs,
              Offset=[0, 0, 0], MaximumSimulationTime=steps_done * 0.005, SaveInterval=0)
    ref = oracle.run(st, wind, seeds_sub, real64=True)
    if ref["steps"] != steps_done:
        return None
    p_ref = ref["trajectories"][:, -1, :]
    both = (ref["status"] == 0)
    d = np.abs(got_sub[both].astype(np.float64) - p_ref[both]) / (2 * math.pi)
    return {"rms": float(np.sqrt(np.mean(d * d))), "max": float(d.max()),
            "particles": int(both.sum())}


def cpu_baseline(wind, gs, seconds_hint=15.0):
    """Plain FP32 CPU path (the oracle's FP32 restatement, all host threads) on a bounded
    sample of the same workload: 65,536 particles over 200 steps (one window)."""
    import oracle
    from pygranite_amd import synthetic
    n, steps = 65536, 200
    seeds = synthetic.uniform_seeds(11, n, [0.0] * 3, [2 * math.pi] * 3)
    st = dict(Space=0, Integrator=1, DeltaT=0.005, DataTimeDistance=1.0, GridScale=gs,
              Offset=[0, 0, 0], MaximumSimulationTime=steps * 0.005, SaveInterval=1,
              CurvatureMode=3, AdditionalCompute=["sqrt(u*u+v*v+w*w)"])
    threads = min(os.cpu_count() or 1, 16)
    t0 = time.perf_counter()
    r = oracle.run(st, wind[:3], seeds, real64=False, nthreads=threads)
    el = time.perf_counter() - t0
    return {"value": n * r["steps"] / el, "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{n} particles x {r['steps']} RK4 steps of the same 3D ABC workload "
                      f"(FP32 oracle restatement, {threads} threads, {el:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--chunk", type=int, default=50)
    ap.add_argument("--particles", type=int, default=1 << 20)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path shards by "
              "particles with no exchange during integration, see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    from pygranite_amd import capi, synthetic
    if capi.device_count() == 0:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    ctx = capi.Context(0)
    info = ctx.info()
    wind, gs = synthetic.abc_flow(args.grid, 16)
    seeds = synthetic.uniform_seeds(1, args.particles, [0.0] * 3, [2 * math.pi] * 3)
    exprs = ["sqrt(u*u+v*v+w*w)"]
    print(f"# device {info['name']} CUs={info['compute_units']}", file=sys.stderr, flush=True)

    res = {}
    # fp32: the plain FP32 GPU form (FP32 slots sampled separately, caller particle order);
    # fp32_opt: FP32 with the locality layout; fp16: FP16 storage with the locality layout
    forms = (("fp32", capi.F32, False, False), ("fp32_opt", capi.F32, True, True),
             ("fp16", capi.F16, True, True))
    for name, dtype, paired, reorder in forms:
        print(f"# form {name}: {args.warmup}+{args.steps} launches of {args.chunk} steps, "
              f"{args.particles} particles", file=sys.stderr, flush=True)
        res[name] = run_form(capi, ctx, wind, seeds, gs, dtype, args.chunk, args.steps,
                             args.warmup, exprs, paired=paired, reorder=reorder)
        r = res[name]
        r["psps"] = args.particles * args.chunk * args.steps / r["wall_s"]
        r["psps_kernel"] = args.particles * args.chunk * args.steps / (r["kernel_ms"] / 1e3)
        print(f"#   {name}: {r['psps']:.3e} particle-steps/s (wall), kernel avg "
              f"{r['kernel_ms'] / r['launches']:.3f} ms/launch", file=sys.stderr, flush=True)

    f16, f32 = res["fp16"], res["fp32"]
    outputs = 3 * 4 + 4 + 4  # saved position, curvature, expression per particle-step
    B16 = algorithmic_bytes(2, outputs)
    B32 = algorithmic_bytes(4, outputs)
    steps_per_launch = f16["steps_done"] / max(1, f16["launches"] + 0)  # informational
    avg_ms = f16["kernel_ms"] / f16["launches"]
    ps_per_launch = args.particles * args.chunk * args.steps / f16["launches"]
    achieved = ps_per_launch * B16 / (avg_ms / 1e3) / 1e9

    err = None
    if not args.no_error:
        sub = 256
        err = fp64_error(wind, seeds[:sub], gs, f16["steps_done"], f16["positions"][:sub])
    cpu = None if args.no_cpu else cpu_baseline(wind, gs)

    line = {
        "metric": "RK4 3D particle-steps/s vs plain FP32 at <=1e-4 RMS err; p50/p99 job latency",
        "value": f16["psps"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f16["wall_s"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (fp16 field storage)",
        "data": "synthetic (analytic ABC flow, SplitMix64 seeds)",
        "config": {"workload": "3D ABC flow RK4, 128^3 x 16 slices, 1,048,576 particles, "
                               "dt 0.005, curvature + sqrt(u*u+v*v+w*w) saved every step",
                   "particles": args.particles, "grid": args.grid,
                   "steps_per_launch": args.chunk, "storage": "fp16"},
        "plain_fp32_value": f32["psps"],
        "fp32_locality_value": res["fp32_opt"]["psps"],
        "speedup_vs_plain_fp32": f16["psps"] / f32["psps"],
        "kernel_psps_fp16": f16["psps_kernel"],
        "kernel_psps_fp32": f32["psps_kernel"],
        "fp64_error": err,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_advance<3,RK4,half>", "bytes_per_particle_step": B16,
                     "avg_launch_ms": avg_ms},
        "cpu_baseline": cpu,
        "device": info["name"],
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
