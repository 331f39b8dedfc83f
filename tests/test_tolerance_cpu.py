"""The exact-weight FP32 oracle against its 8-bit texture-weight emulation (rounded and
truncated weights), within the bound the field fixes (tests/tolerance.py); small sizes on the
CPU. The same comparison runs on the GPU at the full configurations (test_gpu_tolerance.py)."""
import math

import numpy as np
import pytest

import oracle
from pygranite_amd import synthetic
from tolerance import DELTA, node_differences, position_bound

WEIGHTS = {"Quantized8": "q8", "Quantized8Truncated": "q8_trunc"}


def check(st, wind, seeds, steps, h, rk4, label):
    r = oracle.run(st, wind, seeds, real64=False)
    exact = r["trajectories"].astype(np.float64)
    D = node_differences(wind, (0, 1))
    for mode, delta in DELTA.items():
        rq = oracle.run(st, wind, seeds, real64=False, weights=WEIGHTS[mode])
        # particles stopped in either run (the Wind threshold near stagnation points) leave the
        # comparison: a stop is a threshold decision, outside a Lipschitz bound
        live = (r["status"] == 0) & (rq["status"] == 0)
        assert live.mean() > 0.98
        dev = np.abs(rq["trajectories"][live].astype(np.float64) - exact[live]).max(axis=(0, 2))
        bound = position_bound(D, h, st["DeltaT"], steps, delta, rk4, np.abs(exact).max())
        assert dev.shape == bound.shape
        assert np.all(dev <= bound), f"{label} {mode}: step {np.argmax(dev > bound)}"
        # the weights really differ, and the bound is not vacuous at the end
        assert dev[-1] > 0 and dev[-1] > 1e-3 * bound[-1], (label, mode, dev[-1], bound[-1])


def test_abc_rk4_within_weight_bound():
    n, steps = 32, 100
    wind, gs = synthetic.abc_flow(n, 4)
    st = dict(Space=0, Integrator=1, DeltaT=0.005, DataTimeDistance=1.0, GridScale=gs,
              Offset=[0, 0, 0], MaximumSimulationTime=(steps - 0.5) * 0.005, SaveInterval=1)
    lo, hi = 1.8, 2 * math.pi - 1.8  # 100 steps move a particle less than 1.6
    seeds = synthetic.uniform_seeds(5, 512, [lo] * 3, [hi] * 3)
    check(st, wind, seeds, steps, gs, True, "abc")


def test_double_gyre_euler_within_weight_bound():
    steps = 100
    wind, gs = synthetic.double_gyre(128, 64, 3)
    st = dict(Space=1, Integrator=0, DeltaT=0.01, DataTimeDistance=1.0, GridScale=gs,
              Offset=[0, 0], MaximumSimulationTime=(steps - 0.5) * 0.01, SaveInterval=1)
    seeds = synthetic.uniform_seeds(6, 1024, [0.1, 0.1], [1.9, 0.9])
    check(st, wind, seeds, steps, gs, False, "gyre")


def test_bound_catches_a_planted_error():
    """The bound is tight enough to fail a plausible mistake: weights quantised to 1/64
    (four times the truncation error) break it on the ABC flow."""
    n, steps = 32, 100
    wind, gs = synthetic.abc_flow(n, 4)
    st = dict(Space=0, Integrator=1, DeltaT=0.005, DataTimeDistance=1.0, GridScale=gs,
              Offset=[0, 0, 0], MaximumSimulationTime=(steps - 0.5) * 0.005, SaveInterval=1)
    seeds = synthetic.uniform_seeds(5, 512, [1.8] * 3, [2 * math.pi - 1.8] * 3)
    exact = oracle.run(st, wind, seeds, real64=False)["trajectories"].astype(np.float64)
    # planted: the field itself perturbed by 4 x the trunc per-sample bound, in a pattern the
    # weights cannot produce (a constant offset), stands in for a too-coarse weight grid
    D = node_differences(wind, (0, 1))
    eps = 4 * DELTA["Quantized8Truncated"] * D.sum()
    bad = oracle.run(st, (wind + np.float32(eps)).astype(np.float32), seeds, real64=False)["trajectories"]
    dev = np.abs(bad.astype(np.float64) - exact).max(axis=(0, 2))
    bound = position_bound(D, gs, 0.005, steps, DELTA["Quantized8Truncated"], True, np.abs(exact).max())
    assert np.any(dev > bound)
