"""Committed golden fixtures (tests/golden/*.npz, written by make_golden.py from the FP32
oracle): the oracle must keep reproducing them bit for bit (CPU), and the HIP path must match
them bit for bit (GPU)."""
import os
import sys

import numpy as np
import pytest

import oracle
from helpers import run_product

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)
import make_golden  # noqa: E402

NAMES = sorted(make_golden.CASES)


def load(name):
    with np.load(os.path.join(HERE, f"{name}.npz")) as z:
        return {k: z[k] for k in z.files}


def settings_for_product(st):
    names = {0: "Space3D", 1: "Space2D"}
    out = dict(st)
    out["Space"] = names[st["Space"]]
    out["Integrator"] = ["ExplicitEuler", "ClassicRungeKutta"][st["Integrator"]]
    if "CurvatureMode" in st:
        out["CurvatureMode"] = ["Off", "TotalCurvature", "FastTotalCurvature",
                                "IndividualAndTotalCurvature"][st["CurvatureMode"]]
    return out


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    st, wind, seeds = make_golden.inputs(make_golden.CASES[name])
    g = load(name)
    r = oracle.run(st, wind, seeds, real64=False, nthreads=1)
    np.testing.assert_array_equal(r["trajectories"].astype(np.float32), g["trajectories"])
    np.testing.assert_array_equal(r["status"], g["status"])
    np.testing.assert_array_equal(r["abort_step"], g["abort_step"])
    if "curvatures" in g:
        np.testing.assert_array_equal(r["curvatures"].astype(np.float32), g["curvatures"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_golden(pg, gpu_available, name):
    st, wind, seeds = make_golden.inputs(make_golden.CASES[name])
    g = load(name)
    r = run_product(pg, settings_for_product(st), wind, seeds)
    np.testing.assert_array_equal(r["trajectories"], g["trajectories"])
    np.testing.assert_array_equal(r["status"], g["status"])
    np.testing.assert_array_equal(r["abort_step"].astype(np.int64), g["abort_step"])
    if "curvatures" in g:
        np.testing.assert_array_equal(r["curvatures"], g["curvatures"])
        np.testing.assert_array_equal(r["compute"][0], g["compute0"])
