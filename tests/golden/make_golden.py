#!/usr/bin/env python3
"""Writes the small golden fixtures of tests/golden/ from the oracle.

The reference ships no golden vectors for this path, so these fixtures pin the oracle itself
(a change in its arithmetic shows up as a fixture mismatch) and give the GPU tests fixed
expected outputs. Each fixture holds the inputs' recipe (settings + synthetic-field and seed
parameters, regenerated deterministically) and the FP32 oracle outputs.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from pygranite_amd import synthetic  # noqa: E402

CASES = {
    "gyre2d_euler": dict(
        field=("double_gyre", dict(nx=128, ny=64, nslices=4)),
        seeds=dict(seed=0, n=64, lo=[0.05, 0.05], hi=[1.95, 0.95]),
        settings=dict(Space=1, Integrator=0, DeltaT=0.01, DataTimeDistance=1.0,
                      Offset=[0.0, 0.0], MaximumSimulationTime=2.5, SaveInterval=25)),
    "abc3d_rk4_curv_expr": dict(
        field=("abc_flow", dict(n=16, nslices=3)),
        seeds=dict(seed=1, n=48, lo=[0.0] * 3, hi=[2 * math.pi] * 3),
        settings=dict(Space=0, Integrator=1, DeltaT=0.005, DataTimeDistance=1.0,
                      Offset=[0.0, 0.0, 0.0], MaximumSimulationTime=1.5, SaveInterval=20,
                      CurvatureMode=3, AdditionalCompute=["sqrt(u*u+v*v+w*w)"])),
}


def inputs(case):
    kind, kw = case["field"]
    wind, gs = getattr(synthetic, kind)(**kw)
    sd = case["seeds"]
    seeds = synthetic.uniform_seeds(sd["seed"], sd["n"], sd["lo"], sd["hi"])
    st = dict(case["settings"], GridScale=gs)
    return st, wind, seeds


def main():
    for name, case in CASES.items():
        st, wind, seeds = inputs(case)
        r = oracle.run(st, wind, seeds, real64=False, nthreads=1)
        out = {
            "trajectories": r["trajectories"].astype(np.float32),
            "status": r["status"].astype(np.int32),
            "abort_step": r["abort_step"].astype(np.int64),
        }
        if r["curvatures"].size:
            out["curvatures"] = r["curvatures"].astype(np.float32)
            out["total_curvature"] = r["total_curvature"].astype(np.float32)
        for i, c in enumerate(r["compute"]):
            out[f"compute{i}"] = c.astype(np.float32)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        with open(os.path.join(HERE, f"{name}.json"), "w") as f:
            json.dump({k: v for k, v in case.items()}, f, indent=1)
        print(name, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
