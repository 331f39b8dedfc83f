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
