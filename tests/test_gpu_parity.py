"""GPU parity: the HIP path (through pygranite's API and the C ABI) against the CPU oracle.

FP32 runs are compared bit for bit with the oracle's FP32 restatement (same IEEE operations,
no contraction on either side): every saved position, every status, every abort step. FP16 /
BF16 storage is compared bit for bit with the FP32 oracle fed the same rounded slices. Accuracy
against the FP64 oracle is checked as domain-relative error.
"""
import math

import numpy as np
import pytest

import oracle
from helpers import assert_bitexact, domain_rel_error, run_product
from pygranite_amd import synthetic

pytestmark = pytest.mark.gpu


def c0_settings(steps=1000, save=10):
    wind, gs = synthetic.double_gyre(512, 256, 11)
    st = dict(Space="Space2D", Integrator="ExplicitEuler", DeltaT=0.01, DataTimeDistance=1.0,
              GridScale=gs, Offset=[0.0, 0.0], MaximumSimulationTime=steps * 0.01,
              SaveInterval=save, BorderMode="Block")
    return st, wind, gs


def c1_settings(n=32, steps=400):
    wind, gs = synthetic.abc_flow(n, 16)
    st = dict(Space="Space3D", Integrator="ClassicRungeKutta", DeltaT=0.005,
              DataTimeDistance=1.0, GridScale=gs, Offset=[0.0, 0.0, 0.0],
              MaximumSimulationTime=steps * 0.005, SaveInterval=1,
              CurvatureMode="IndividualAndTotalCurvature",
              AdditionalCompute=["sqrt(u*u+v*v+w*w)"])
    return st, wind, gs


def test_c0_double_gyre_euler_bitexact(pg, gpu_available):
    st, wind, gs = c0_settings()
    h = np.array(gs)
    seeds = synthetic.uniform_seeds(0, 4096, h, np.array([2.0, 1.0]) - h)
    gpu = run_product(pg, st, wind, seeds)
    ref = oracle.run(st, wind, seeds, real64=False)
    assert gpu["trajectories"].shape == (4096, 101, 2)
    assert_bitexact(gpu, ref, label="C0 ")
    # the FP32 path is close to the FP64 oracle
    ref64 = oracle.run(st, wind, seeds, real64=True)
    mx, rms = domain_rel_error(gpu["trajectories"], ref64["trajectories"], [2.0, 1.0])
    assert rms < 1e-5, (mx, rms)


def test_c1_abc_rk4_curvature_expr_bitexact(pg, gpu_available):
    st, wind, gs = c1_settings()
    seeds = synthetic.uniform_seeds(1, 2048, [0.0] * 3, [2 * math.pi] * 3)
    gpu = run_product(pg, st, wind, seeds)
    ref = oracle.run(st, wind, seeds, real64=False)
    assert_bitexact(gpu, ref, fields=("trajectories", "curvatures", "total_curvature"),
                    label="C1 ")
    # sqrt is correctly rounded on both sides -> the expression is bit-exact as well
    np.testing.assert_array_equal(gpu["compute"][0], ref["compute"][0].astype(np.float32))
    assert gpu["curvatures"].shape == (2048, 399)
    assert gpu["compute"][0].shape == (2048, 400)


@pytest.mark.parametrize("storage,rounder", [("Float16", synthetic.round_fp16),
                                             ("BFloat16", synthetic.round_bf16)])
def test_c1_half_storage_bitexact(pg, gpu_available, storage, rounder):
    st, wind, gs = c1_settings(n=32, steps=200)
    st = dict(st, WindfieldStorage=storage)
    seeds = synthetic.uniform_seeds(2, 1024, [0.0] * 3, [2 * math.pi] * 3)
    gpu = run_product(pg, st, wind, seeds)
    ref = oracle.run(st, rounder(wind), seeds, real64=False)
    assert_bitexact(gpu, ref, fields=("trajectories", "curvatures"), label=storage + " ")


def test_texture_weight_emulation_bitexact(pg, gpu_available):
    st, wind, gs = c0_settings(steps=300, save=5)
    st = dict(st, TextureWeights="Quantized8")
    seeds = synthetic.uniform_seeds(3, 2048, gs, np.array([2.0, 1.0]) - np.array(gs))
    gpu = run_product(pg, st, wind, seeds)
    ref = oracle.run(st, wind, seeds, real64=False, weights="q8")
    assert_bitexact(gpu, ref, label="q8 ")


def test_precision_form_kahan_time64_bitexact(pg, gpu_available):
    st, wind, gs = c1_settings(n=32, steps=300)
    st = dict(st, KahanSummation=True, TimeFP64=True, WindfieldStorage="Float16")
    seeds = synthetic.uniform_seeds(4, 1024, [0.0] * 3, [2 * math.pi] * 3)
    gpu = run_product(pg, st, wind, seeds)
    ref = oracle.run(st, synthetic.round_fp16(wind), seeds, real64=False, time64=True, kahan=True)
    assert_bitexact(gpu, ref, label="kahan ")


def test_stretched_z_precision_form_bitexact(pg, gpu_available):
    """Stretched z axis (ZLevels extension) on an atmosphere-like field: FP32, FP16 and BF16 storage,
    RK4, Kahan positions, FP64 time, the expression w*1000/(abs(u)+abs(v)+1) saved every step,
    bit for bit against the FP32 oracle on the same levels (and the same rounded slices)."""
    nx = ny = 48
    nz, S = 20, 4
    wind, gs, zlev = synthetic.atmosphere(nx, ny, nz, S, seed=7, dx=3000.0, dz0=20.0, ratio=1.08)
    dom = [nx * gs[0], ny * gs[1], float(zlev[-1])]
    n = 2048
    xy = synthetic.uniform_seeds(23, n, [0.0, 0.0], [dom[0], dom[1]])
    z = synthetic.loguniform_heights(24, n, float(zlev[1]), float(zlev[-1]))
    seeds = np.ascontiguousarray(np.concatenate([xy, z[:, None]], axis=1), np.float32)
    for storage, rounder in (("Float32", lambda a: a), ("Float16", synthetic.round_fp16),
                             ("BFloat16", synthetic.round_bf16)):
        st = dict(Space="Space3D", Integrator="ClassicRungeKutta", DeltaT=60.0,
                  DataTimeDistance=3600.0, GridScale=list(gs), Offset=[0.0, 0.0, 0.0],
                  MaximumSimulationTime=150 * 60.0, SaveInterval=1, KahanSummation=True,
                  TimeFP64=True, WindfieldStorage=storage, AllowAbs=True,
                  AdditionalCompute=["w*1000/(abs(u)+abs(v)+1)"], ZLevels=[float(v) for v in zlev])
        gpu = run_product(pg, st, wind, seeds)
        ref = oracle.run(st, rounder(wind), seeds, real64=False, time64=True, kahan=True,
                         zlevels=zlev, allow_abs=True)
        assert_bitexact(gpu, ref, label=f"stretched z {storage} ")
        np.testing.assert_array_equal(gpu["compute"][0], ref["compute"][0].astype(np.float32))
        # a mix of particles alive at the end and stopped on the way (domain exits, stalls)
        assert (ref["status"] != 0).sum() > 100 and (ref["status"] == 0).sum() > 100


@pytest.mark.parametrize("save_interval", [0, 1])
def test_regrouping_inside_windows_changes_nothing(pg, gpu_available, save_interval):
    """The drop-in path regroups particles by cell at window edges and, in runs that save no
    columns, every RegroupInterval steps inside a window (here 16, whatever the live count):
    every output equals a run that never regroups, and both equal the FP32 oracle on a
    subsample."""
    st, wind, gs = c1_settings(steps=300)
    st = dict(st, SaveInterval=save_interval)
    seeds = synthetic.uniform_seeds(21, 6000, [0.0] * 3, [2 * math.pi] * 3)
    regrouped = run_product(pg, dict(st, RegroupInterval=16, RegroupMinLive=0), wind, seeds)
    plain = run_product(pg, dict(st, SortParticles=False), wind, seeds)
    for f in ("trajectories", "curvatures", "total_curvature", "abort_step"):
        np.testing.assert_array_equal(regrouped[f], plain[f], err_msg=f)
    np.testing.assert_array_equal(regrouped["status"], plain["status"])
    np.testing.assert_array_equal(regrouped["compute"][0], plain["compute"][0])
    assert (plain["status"] != 0).sum() > 100  # particles stopped at different steps
    ref = oracle.run(st, wind, seeds[:64], real64=False)
    sub = {k: (v[:64] if isinstance(v, np.ndarray) else v) for k, v in regrouped.items()}
    assert_bitexact(sub, ref, fields=("trajectories", "curvatures") if save_interval else ("trajectories",),
                    label="regrouped ")


@pytest.mark.parametrize("expr", ["sqrt(u*u+v*v+w*w)", "u*u+v*v+w*w", "x*u+y*v+z*w", "sqrt(v*v+u*u+w*w)",
                                  "u*u+v*v"])
def test_single_dot_programs_bitexact(pg, gpu_available, expr):
    """Programs of one DOT instruction skip the interpreter's dispatch loop (the squared speed
    u*u + v*v [+ w*w] reads the velocity registers directly): every saved value equals the FP32
    oracle's evaluation bit for bit, as the interpreted programs do."""
    st, wind, gs = c1_settings(steps=120)
    st = dict(st, AdditionalCompute=[expr])
    seeds = synthetic.uniform_seeds(23, 3000, [0.0] * 3, [2 * math.pi] * 3)
    gpu = run_product(pg, st, wind, seeds)
    ref = oracle.run(st, wind, seeds, real64=False)
    assert_bitexact(gpu, ref, fields=("trajectories",), label=expr + " ")
    np.testing.assert_array_equal(gpu["compute"][0], ref["compute"][0].astype(np.float32))
