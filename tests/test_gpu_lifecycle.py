"""Object lifetimes through the C ABI: a context with live fields or jobs is not destroyed, and
the Python Context closes its children before itself."""
import numpy as np
import pytest

from pygranite_amd import capi

pytestmark = pytest.mark.gpu


def test_context_outlives_children(gpu_available):
    ctx = capi.Context(0)
    f = capi.Field(ctx, 8, 8, 8, 3, capi.F32, nslots=2)
    d = capi.JobDesc()
    d.dim, d.integrator, d.dt, d.dtd = 3, 1, 0.1, 1.0
    for k in range(3):
        d.grid_scale[k], d.domain[k] = 1.0, 8.0
    j = capi.Job(ctx, d, np.ones((4, 3), np.float32))
    # the raw destroy refuses while children live
    assert capi.lib().gr_ctx_destroy(ctx.h) == -5
    assert b"still alive" in capi.lib().gr_last_error()
    # closing the wrapper closes the job and the field first
    ctx.close()
    assert not j.h and not f.h and not ctx.h
    # their own close / __del__ are now no-ops
    j.close()
    f.close()


def test_integrators_on_two_threads(pg, gpu_available):
    """Two Python threads run TrajectoryIntegrator.compute() at the same time (each integrator
    owns its context; the GIL is released around device work): each result equals the same
    integration run alone."""
    import math
    import threading

    from helpers import run_product
    from pygranite_amd import synthetic

    wind, gs = synthetic.abc_flow(24, 4)
    st = dict(Space="Space3D", Integrator="ClassicRungeKutta", DeltaT=0.01, DataTimeDistance=1.0,
              GridScale=gs, Offset=[0.0, 0.0, 0.0], MaximumSimulationTime=2.5, SaveInterval=5,
              CurvatureMode="IndividualAndTotalCurvature", AdditionalCompute=["sqrt(u*u+v*v+w*w)"])
    seeds = [synthetic.uniform_seeds(60 + k, 6000, [0.0] * 3, [2 * math.pi] * 3) for k in range(2)]
    alone = [run_product(pg, st, wind, s) for s in seeds]
    got, errs = [None, None], []

    def work(k):
        try:
            for _ in range(3):
                got[k] = run_product(pg, st, wind, seeds[k])
        except Exception as e:  # reported below
            errs.append(e)

    th = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errs, errs
    for k in range(2):
        for f in ("trajectories", "status", "abort_step", "curvatures", "total_curvature"):
            np.testing.assert_array_equal(got[k][f], alone[k][f], err_msg=f)
        np.testing.assert_array_equal(got[k]["compute"][0], alone[k]["compute"][0])


def test_copy_bandwidth_measures_the_device(gpu_available):
    """The streaming-copy measurement returns a device-to-device rate between 1 TB/s and the
    published 8 TB/s peak plus margin, and refuses sizes that are not whole 16-byte units."""
    from pygranite_amd import capi
    ctx = capi.Context(0)
    gbs = ctx.copy_bandwidth(1 << 28, 3)
    assert 1000.0 < gbs < 10000.0, gbs
    import pytest
    with pytest.raises(capi.GraniteError):
        ctx.copy_bandwidth(100, 1)
    ctx.close()
