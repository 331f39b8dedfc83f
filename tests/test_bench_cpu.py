"""The measurement script's host logic, on the CPU: the one-GPU guard, the algorithmic bytes the
roofline divides, and the classification of termination mismatches against the FP64 oracle."""
import os
import subprocess
import sys

import numpy as np

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_algorithmic_bytes_per_particle_step():
    # 4 stages x 2 slices x 3 components x 8 corners at the storage width, position read and
    # write plus status, and the saved outputs (position 12 B, curvature 4 B, expression 4 B)
    assert bench.algorithmic_bytes(4, 20) == 4 * 2 * 3 * 8 * 4 + 28 + 20 == 816
    assert bench.algorithmic_bytes(2, 20) == 4 * 2 * 3 * 8 * 2 + 28 + 20 == 432


def _paths(r64, r32):
    return lambda idx: ({k: np.asarray(v)[idx] for k, v in r64.items()},
                        {k: np.asarray(v)[idx] for k, v in r32.items()})


def test_termination_classification_and_error():
    ext = [8.0, 4.0, 2.0]
    ref = np.array([[1.0, 1.0, 1.0], [2.0, 2.0, 1.0], [3.0, 1.0, 1.0], [4.0, 2.0, 1.0], [5.0, 1.0, 1.0]])
    pos = ref.copy()
    pos[0] += [0.08, 0.0, 0.0]   # 1 % of the x extent
    pos[1] += [0.0, 0.0, -0.04]  # 2 % of the z extent
    st_ref = np.array([0, 0, 3, 3, 3], np.int32)
    ab_ref = np.array([-1, -1, 10, 10, 10])
    st = np.array([0, 0, 3, 3, 4], np.int32)
    ab = np.array([-1, -1, 12, 11, 10])
    # step-by-step paths of the three mismatched particles (2, 3, 4): 13 columns, unit steps in x
    T = np.zeros((3, 13, 3))
    T[:, :, 0] = np.arange(13)[None, :]
    T32 = T.copy()
    T32[0, 10, 0] += 5.0   # particle 2: 5 step lengths apart at the first stop -> divergence
    T32[1, 10, 0] += 0.5   # particle 3: half a step apart, same reason -> threshold
    T32[2, 10, 0] += 0.25  # particle 4: a quarter step apart, reasons differ -> other
    r64 = {"trajectories": T, "abort_step": np.array([10, 10, 10]), "status": np.array([3, 3, 3])}
    r32 = {"trajectories": T32, "abort_step": np.array([12, 11, 10]), "status": np.array([3, 3, 4])}
    seen = []

    def paths(idx):
        seen.append(list(idx))
        return r64, r32

    out = bench.termination_and_error(pos, st, ab, ref, st_ref, ab_ref, ext, paths)
    assert seen == [[2, 3, 4]]
    assert out["termination_agree"] == 2 and out["termination_mismatch"] == 3
    assert (out["mismatch_divergence"], out["mismatch_threshold"], out["mismatch_other"]) == (1, 1, 1)
    # errors over the two agreeing particles, each coordinate divided by its extent
    d = np.array([[0.01, 0.0, 0.0], [0.0, 0.0, 0.02]])
    assert np.isclose(out["max"], 0.02)
    assert np.isclose(out["rms"], np.sqrt(np.mean(d * d)))
    assert out["alive_at_end"] == 2
    assert np.isclose(out["median_separation_in_steps"], 0.5)
