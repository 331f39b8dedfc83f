"""GPU parity at the configurations' full sizes.

* The 2D double-gyre configuration in full: 512 x 256 x 11 slices, 65,536 particles, Euler,
  1000 steps saved every 10th, bit for bit against the FP32 oracle.
* The 3D ABC configuration on its real 128^3 x 16-slice field over the whole 2000 steps (RK4,
  per-step curvature and sqrt(u*u+v*v+w*w) saved every step), on a 4,096-particle subsample of
  its seeds: the drop-in API (time-paired layout, particles regrouped by cell), the plain FP32
  form in the batched kernel (slots sampled separately, caller order), and the FP16 batched form
  against the FP32 oracle fed the same FP16-rounded slices -- all bit for bit.
"""
import math

import numpy as np
import pytest

import oracle
from helpers import assert_bitexact, run_product
from pygranite_amd import capi, synthetic

pytestmark = pytest.mark.gpu


def test_double_gyre_full_configuration_bitexact(pg, gpu_available):
    wind, gs = synthetic.double_gyre(512, 256, 11)
    seeds = synthetic.uniform_seeds(0, 65536, [gs[0], gs[1]], [2.0 - gs[0], 1.0 - gs[1]])
    st = dict(Space="Space2D", Integrator="ExplicitEuler", DeltaT=0.01, DataTimeDistance=1.0,
              GridScale=gs, Offset=[0.0, 0.0], MaximumSimulationTime=10.0, SaveInterval=10)
    gpu = run_product(pg, st, wind, seeds)
    ref = oracle.run(st, wind, seeds, real64=False, nthreads=16)
    assert gpu["trajectories"].shape == (65536, 101, 2)
    assert_bitexact(gpu, ref, label="double gyre full ")


def abc_full():
    wind, gs = synthetic.abc_flow(128, 16)
    seeds = synthetic.uniform_seeds(1, 1 << 20, [0.0] * 3, [2 * math.pi] * 3)
    # every 256th seed of the configuration's 2^20: spread over the whole domain
    return wind, gs, np.ascontiguousarray(seeds[::256])


ABC_ST = dict(Space="Space3D", Integrator="ClassicRungeKutta", DeltaT=0.005, DataTimeDistance=1.0,
              Offset=[0.0, 0.0, 0.0], MaximumSimulationTime=10.0, SaveInterval=1,
              CurvatureMode="IndividualAndTotalCurvature", AdditionalCompute=["sqrt(u*u+v*v+w*w)"])


@pytest.fixture(scope="module")
def abc_ref():
    wind, gs, seeds = abc_full()
    st = dict(ABC_ST, GridScale=gs)
    ref = oracle.run(st, wind, seeds, real64=False, nthreads=16)
    assert ref["steps"] == 2000 and seeds.shape[0] == 4096
    return wind, gs, seeds, st, ref


def test_abc_full_field_2000_steps_drop_in_bitexact(pg, gpu_available, abc_ref):
    wind, gs, seeds, st, ref = abc_ref
    gpu = run_product(pg, dict(st, RegroupMinLive=0, RegroupInterval=64), wind, seeds)
    assert gpu["trajectories"].shape == (4096, 2001, 3)
    assert_bitexact(gpu, ref, fields=("trajectories", "curvatures", "total_curvature"),
                    label="ABC full drop-in ")
    np.testing.assert_array_equal(gpu["compute"][0], ref["compute"][0].astype(np.float32))


def _batched(wind, gs, seeds, dtype, paired, reorder):
    ctx = capi.Context(0)
    try:
        f = capi.Field(ctx, 128, 128, 128, 3, dtype=dtype, nslots=16, address=0)
        for s in range(16):
            f.upload(s, [wind[s, c] for c in range(3)])
        d = capi.JobDesc()
        d.dim, d.integrator, d.curvature, d.windfield_mode = 3, 1, 3, 1
        d.dt, d.dtd = 0.005, 1.0
        for k in range(3):
            d.grid_scale[k] = gs[k]
            d.domain[k] = float(np.float32(128) * np.float32(gs[k]))
        d.save_interval = 1
        d.nprog = 1
        d.prog[0] = capi.compile_expr("sqrt(u*u+v*v+w*w)", dim=3, curvature=3)
        parts = np.array_split(np.arange(seeds.shape[0]), 4)
        jobs = [capi.Job(ctx, d, np.ascontiguousarray(seeds[p]), save_capacity=2000) for p in parts]
        for j in jobs:
            j.set_option(capi.OPT_PAIRED, paired)
        if reorder:
            ctx.reorder(jobs)
        ends, _ = ctx.batch_run([(j, f, dict(max_steps=2000, max_sim_time=10.0)) for j in jobs],
                                ends=True)
        assert all(e["steps"] == 2000 for e in ends)
        pos = np.concatenate([np.transpose(j.read_saved(capi.OUT_POSITIONS, 0, 2000), (1, 0, 2))
                              for j in jobs])
        curv = np.concatenate([j.read_saved(capi.OUT_CURVATURE, 0, 2000).T for j in jobs])
        expr = np.concatenate([j.read_saved(capi.OUT_EXPR0, 0, 2000).T for j in jobs])
        status = np.concatenate([j.read(capi.OUT_STATUS) for j in jobs])
        ab = np.concatenate([j.read(capi.OUT_ABORT_STEP) for j in jobs])
        total = np.concatenate([j.read(capi.OUT_TOTAL_CURVATURE) for j in jobs])
        for j in jobs:
            j.close()
        f.close()
    finally:
        ctx.close()
    return {"trajectories": pos, "curvatures": curv[:, 1:], "compute": [expr], "status": status,
            "abort_step": ab, "total_curvature": total}


def test_abc_full_field_plain_fp32_batched_bitexact(gpu_available, abc_ref):
    wind, gs, seeds, st, ref = abc_ref
    gpu = _batched(wind, gs, seeds, capi.F32, paired=0, reorder=False)
    want = {"trajectories": ref["trajectories"][:, 1:], "curvatures": ref["curvatures"],
            "status": ref["status"], "abort_step": ref["abort_step"],
            "total_curvature": ref["total_curvature"]}
    assert_bitexact(gpu, want, fields=("trajectories", "curvatures", "total_curvature"),
                    label="ABC full plain FP32 ")
    np.testing.assert_array_equal(gpu["compute"][0], ref["compute"][0].astype(np.float32))


def test_abc_full_field_fp16_batched_bitexact(gpu_available, abc_ref):
    wind, gs, seeds, st, _ = abc_ref
    ref = oracle.run(st, synthetic.round_fp16(wind), seeds, real64=False, nthreads=16)
    gpu = _batched(wind, gs, seeds, capi.F16, paired=1, reorder=True)
    want = {"trajectories": ref["trajectories"][:, 1:], "curvatures": ref["curvatures"],
            "status": ref["status"], "abort_step": ref["abort_step"],
            "total_curvature": ref["total_curvature"]}
    assert_bitexact(gpu, want, fields=("trajectories", "curvatures", "total_curvature"),
                    label="ABC full FP16 batched ")
    np.testing.assert_array_equal(gpu["compute"][0], ref["compute"][0].astype(np.float32))
