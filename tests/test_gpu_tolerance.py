"""The HIP path with exact FP32 weights against the same path with the reference's 8-bit
texture-weight emulation (rounded and truncated), at the full double-gyre (512 x 256, Euler,
dt 0.01) and ABC (128^3, RK4, dt 0.005) fields: every saved position stays within the bound the
field fixes (tests/tolerance.py), step by step, over 16,384 interior particles and 100 steps
(the horizon in which no particle reaches the domain edge). Particles that stop in either run
(the Wind threshold |dp|^2 < 1e-9, met only near the double gyre's stagnation points) leave the
comparison: a stop is a threshold decision, outside what a Lipschitz bound speaks of."""
import math

import numpy as np
import pytest

from helpers import run_product
from pygranite_amd import synthetic
from tolerance import DELTA, node_differences, position_bound

pytestmark = pytest.mark.gpu


def check(pg, st, wind, seeds, steps, h, rk4, label):
    exact = run_product(pg, st, wind, seeds)
    ex = np.asarray(exact["trajectories"], np.float64)
    D = node_differences(wind, (0, 1))
    for mode, delta in DELTA.items():
        q = run_product(pg, dict(st, TextureWeights=mode), wind, seeds)
        live = (exact["status"] == 0) & (q["status"] == 0)
        assert live.mean() > 0.99, (label, mode, int((~live).sum()))
        dev = np.abs(np.asarray(q["trajectories"], np.float64)[live] - ex[live]).max(axis=(0, 2))
        bound = position_bound(D, h, st["DeltaT"], steps, delta, rk4, np.abs(ex).max())
        assert dev.shape == bound.shape == (steps + 1,)
        assert np.all(dev <= bound), f"{label} {mode}: first step over the bound {np.argmax(dev > bound)}"
        assert dev[-1] > 1e-3 * bound[-1], (label, mode, dev[-1], bound[-1])
        print(f"{label} {mode}: max deviation {dev[-1]:.3e} <= bound {bound[-1]:.3e} after {steps} steps")


def test_double_gyre_full_field_within_weight_bound(pg, gpu_available):
    steps = 100
    wind, gs = synthetic.double_gyre(512, 256, 11)
    st = dict(Space="Space2D", Integrator="ExplicitEuler", DeltaT=0.01, DataTimeDistance=1.0,
              GridScale=gs, Offset=[0.0, 0.0], MaximumSimulationTime=(steps - 0.5) * 0.01,
              SaveInterval=1, BorderMode="Block")
    seeds = synthetic.uniform_seeds(31, 16384, [0.1, 0.1], [1.9, 0.9])
    check(pg, st, wind, seeds, steps, gs, False, "double gyre")


def test_abc_full_field_within_weight_bound(pg, gpu_available):
    steps = 100
    wind, gs = synthetic.abc_flow(128, 16)
    st = dict(Space="Space3D", Integrator="ClassicRungeKutta", DeltaT=0.005, DataTimeDistance=1.0,
              GridScale=gs, Offset=[0.0, 0.0, 0.0], MaximumSimulationTime=(steps - 0.5) * 0.005,
              SaveInterval=1)
    seeds = synthetic.uniform_seeds(32, 16384, [1.8] * 3, [2 * math.pi - 1.8] * 3)
    check(pg, st, wind, seeds, steps, gs, True, "ABC")
