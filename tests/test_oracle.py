"""The oracle pinned against closed forms, brute force and the reference's host arithmetic.

The reference ships no tests, fixtures or golden vectors for this path, so every oracle function
is pinned here by mathematics: exact reproduction of (multi)linear fields, closed-form Euler /
RK4 amplification factors, the circle curvature, textbook operator precedence, and the FP32
time arithmetic of the reference's host loop recomputed with numpy.float32.
"""
import math

import numpy as np
import pytest

import oracle
from pygranite_amd import synthetic


def box3(nx=8, ny=8, nz=8, S=3):
    return np.zeros((S, 3, nz, ny, nx), np.float32)


def base3(**kw):
    d = dict(Space=0, Integrator=1, DeltaT=0.1, DataTimeDistance=1.0, GridScale=[1, 1, 1],
             Offset=[0, 0, 0], MaximumSimulationTime=1.0, SaveInterval=1)
    d.update(kw)
    return d


def test_uniform_wind_translation_exact():
    w = box3()
    w[:, 0], w[:, 1], w[:, 2] = 1.0, 0.5, 0.25
    seeds = np.array([[1, 1, 1], [2, 3, 1.5]], np.float32)
    for integ in (0, 1):
        r = oracle.run(base3(Integrator=integ), w, seeds)
        T = r["trajectories"]
        assert T.shape == (2, 11, 3)
        steps = np.arange(11)[None, :, None] * 0.1
        expect = seeds[:, None, :].astype(np.float64) + steps * np.array([1.0, 0.5, 0.25])
        np.testing.assert_allclose(T, expect, rtol=0, atol=1e-6)


@pytest.mark.parametrize("integ", [0, 1])
def test_linear_field_amplification(integ):
    """u = lam * x (exactly trilinear): x_n = x0 * g^n with the Euler / RK4 growth factor."""
    lam, dt, n = 0.5, 0.05, 40
    nx = 64
    h = 1.0
    w = np.zeros((3, 3, 4, 4, nx), np.float32)
    x = np.arange(nx, dtype=np.float64) * h
    w[:, 0] = (lam * x)[None, None, None, :]
    seeds = np.array([[1.0, 1.0, 1.0]], np.float32)
    st = base3(Integrator=integ, DeltaT=dt, MaximumSimulationTime=n * dt, GridScale=[h, h, h],
               DataTimeDistance=100.0)
    r = oracle.run(st, w, seeds)
    z = lam * float(np.float32(dt))  # DeltaT is a float setting
    g = 1 + z if integ == 0 else 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    xs = r["trajectories"][0, :, 0]
    expect = 1.0 * g ** np.arange(xs.size)
    np.testing.assert_allclose(xs, expect, rtol=1e-12)


def test_solid_body_rotation_rk4_closed_form():
    """u = -w*y, v = w*x is bilinear, so sampling is exact; RK4 on it is the linear map with the
    truncated exponential exp(A*dt) (|z|^k / k! terms up to 4)."""
    om, dt, n = 1.0, 0.05, 60
    N = 64
    c = 32.0  # rotation about the grid centre (in index units, h = 1)
    x = np.arange(N, dtype=np.float64)
    X, Y = np.meshgrid(x, x)
    w = np.zeros((3, 2, N, N), np.float32)
    w[:, 0] = -om * (Y - c)
    w[:, 1] = om * (X - c)
    seeds = np.array([[c + 10.0, c]], np.float32)
    st = dict(Space=1, Integrator=1, DeltaT=dt, DataTimeDistance=100.0, GridScale=[1.0, 1.0],
              Offset=[0.0, 0.0], MaximumSimulationTime=n * dt, SaveInterval=1)
    r = oracle.run(st, w, seeds)
    A = np.array([[0.0, -om], [om, 0.0]]) * float(np.float32(dt))
    M = np.eye(2) + A + A @ A / 2 + A @ A @ A / 6 + A @ A @ A @ A / 24
    p = np.array([10.0, 0.0])
    for k in range(n + 1):
        np.testing.assert_allclose(r["trajectories"][0, k] - c, p, atol=1e-9)
        p = M @ p


def test_time_interpolation_linear_in_time():
    """Linear-in-time uniform wind u(t) = a + b t is reproduced exactly by the time lerp."""
    a, b = 0.5, 0.25
    S = 4
    w = np.zeros((S, 3, 4, 4, 40), np.float32)
    for s in range(S):
        w[s, 0] = a + b * s
    seeds = np.array([[2.0, 1.5, 1.5]], np.float32)
    dt = 0.125  # exact in binary, so the FP32 host time is exact
    st = base3(Integrator=1, DeltaT=dt, MaximumSimulationTime=2.0, DataTimeDistance=1.0)
    r = oracle.run(st, w, seeds)
    t = np.arange(r["trajectories"].shape[1]) * dt
    np.testing.assert_allclose(r["trajectories"][0, :, 0], 2.0 + a * t + b * t * t / 2, atol=1e-12)


def brute_sample3(g, x, y, z):
    """Trilinear sample with border-zero addressing by explicit summation of 8 weighted taps."""
    nz, ny, nx = g.shape
    i, j, k = math.floor(x), math.floor(y), math.floor(z)
    fx, fy, fz = x - i, y - j, z - k
    s = 0.0
    for dk in (0, 1):
        for dj in (0, 1):
            for di in (0, 1):
                a, b, c = i + di, j + dj, k + dk
                v = g[c, b, a] if (0 <= a < nx and 0 <= b < ny and 0 <= c < nz) else 0.0
                s += v * (fx if di else 1 - fx) * (fy if dj else 1 - fy) * (fz if dk else 1 - fz)
    return s


def test_border_addressing_brute_force():
    """One Euler step from points near and beyond the upper grid edge: the sampled velocity
    matches explicit weighted taps with zero outside the grid (3D texture border mode)."""
    rng = np.random.default_rng(5)
    nx = ny = nz = 3
    g = rng.uniform(0.5, 1.5, (nz, ny, nx))
    w = np.zeros((2, 3, nz, ny, nx), np.float32)
    w[:, 0] = g
    w[:, 1] = 1.0  # keep ||dp|| above the wind threshold
    pts = np.array([[0.3, 0.4, 0.2], [1.7, 1.2, 0.9], [2.5, 2.5, 2.5], [2.9, 0.1, 1.5]],
                   np.float32)
    dt = 1e-3
    st = dict(Space=0, Integrator=0, DeltaT=dt, DataTimeDistance=1.0, GridScale=[1, 1, 1],
              Offset=[0, 0, 0], MaximumSimulationTime=dt, SaveInterval=1)
    r = oracle.run(st, w, pts)
    g32 = g.astype(np.float32).astype(np.float64)
    for p, traj in zip(pts.astype(np.float64), r["trajectories"]):
        u = brute_sample3(g32, *p)
        assert traj[1, 0] - p[0] == pytest.approx(u * float(np.float32(dt)), rel=1e-9, abs=1e-15)


def test_clamp_addressing_2d():
    """2D uses clamp-to-edge: beyond the last node the edge value is held."""
    N = 4
    w = np.zeros((2, 2, N, N), np.float32)
    w[:, 0] = np.array([1.0, 2.0, 3.0, 4.0], np.float32)[None, None, :]
    w[:, 1] = 0.0
    seeds = np.array([[3.5, 1.0], [0.25, 2.0]], np.float32)  # x in the last cell / first cell
    dt = 1e-3
    st = dict(Space=1, Integrator=0, DeltaT=dt, DataTimeDistance=1.0, GridScale=[1, 1],
              Offset=[0, 0], MaximumSimulationTime=dt, SaveInterval=1)
    r = oracle.run(st, w, seeds)
    dx = r["trajectories"][:, 1, 0] - seeds[:, 0]
    np.testing.assert_allclose(dx / float(np.float32(dt)), [4.0, 1.25], rtol=1e-9)


def test_abort_rules_brute_force():
    """Hand-derived abort reasons and steps: zero wind -> Wind at step 0; leaving the domain
    (N * h, inclusive edge) -> Domain at the exact step; aborted particles freeze in place."""
    N = 8
    w = np.zeros((3, 3, N, N, N), np.float32)
    w[:, 0] = 1.0
    seeds = np.array([[7.45, 4.0, 4.0],   # x reaches 8.0 + > domain after steps
                      [1.0, 1.0, 1.0]], np.float32)
    dt = 0.25
    st = base3(Integrator=0, DeltaT=dt, MaximumSimulationTime=1.0, DataTimeDistance=10.0)
    r = oracle.run(st, w, seeds)
    # particle 0: x = 7.45 + 0.25 k; but border blending halves u in the last cell:
    # sample at x in [7, 8): u = 1 - frac(x) (tap at index 8 reads 0)
    assert r["status"][1] == 0
    zero = np.zeros((2, 3, N, N, N), np.float32)
    r0 = oracle.run(st, zero, seeds)
    assert list(r0["status"]) == [4, 4] and list(r0["abort_step"]) == [0, 0]
    np.testing.assert_array_equal(r0["trajectories"][:, -1], seeds)
    # domain exit: uniform wind on a 2D clamp grid leaves through x = N*h
    w2 = np.zeros((2, 2, N, N), np.float32)
    w2[:, 0] = 1.0
    s2 = np.array([[7.0, 3.0]], np.float32)
    st2 = dict(Space=1, Integrator=0, DeltaT=0.25, DataTimeDistance=10.0, GridScale=[1, 1],
               Offset=[0, 0], MaximumSimulationTime=2.0, SaveInterval=1)
    r2 = oracle.run(st2, w2, s2)
    # x = 7.25, 7.5, 7.75, 8.0 (== Domain: still inside), 8.25 -> Domain at step 4
    assert r2["status"][0] == 2 and r2["abort_step"][0] == 4
    np.testing.assert_allclose(r2["trajectories"][0, -1], [8.0, 3.0])


def test_reference_window_schedule_fp32():
    """Steps per window under the reference's FP32 host time (tpp:973-994): recomputed with
    numpy.float32 and compared with the oracle's schedule."""
    def steps_per_window(dt, dtd, total):
        out = []
        t = np.float32(0)
        counter = 0
        mwt = np.float32(dtd)
        maxs = np.float32(np.float32(total) / np.float32(dt))
        while np.float32(counter) < maxs:
            k = 0
            while np.float32(counter) < maxs:
                k += 1
                counter += 1
                t = np.float32(t + np.float32(dt))
                if t >= mwt:
                    break
            out.append(k)
            mwt = np.float32(mwt + np.float32(dtd))
        return out

    w = box3(4, 4, 4, S=12)
    w[:, 0] = 0.1
    seeds = np.array([[1.0, 1.0, 1.0]], np.float32)
    for dt, total in ((0.01, 10.0), (0.005, 5.0)):
        st = base3(DeltaT=dt, MaximumSimulationTime=total, SaveInterval=1, Integrator=0)
        r = oracle.run(st, w, seeds)
        spw = steps_per_window(dt, 1.0, total)
        assert r["steps"] == sum(spw)
        assert spw[0] in (101, 201)
    # hand-derived: float32(0.01) = 0.0099999998; FP32 time reaches 1.0 after 101 steps and
    # its rounding drifts the later edges to 100, 100, 100, 99 (dt = 0.005: 201, 200, 200, ...)
    assert steps_per_window(0.01, 1.0, 10.0)[:5] == [101, 100, 100, 100, 99]
    assert sum(steps_per_window(0.01, 1.0, 10.0)) == 1000
    assert steps_per_window(0.005, 1.0, 5.0)[0] == 201


def test_slices_limit_windows():
    """S slices give S - 1 windows; the run stops when the loader is exhausted."""
    w = box3(4, 4, 4, S=3)
    w[:, 0] = 0.1
    st = base3(DeltaT=0.25, MaximumSimulationTime=100.0, DataTimeDistance=1.0, Integrator=0)
    r = oracle.run(st, w, np.array([[1.0, 1.0, 1.0]], np.float32))
    assert r["steps"] == 8  # 2 windows x 4 steps


def test_save_interval_columns():
    w = box3(4, 4, 4, S=3)
    w[:, 0] = 0.1
    seeds = np.array([[1.0, 1.0, 1.0]], np.float32)
    for si, cols in ((1, 9), (3, 4), (0, 2)):
        st = base3(DeltaT=0.25, MaximumSimulationTime=2.0, SaveInterval=si, Integrator=0)
        r = oracle.run(st, w, seeds)
        assert r["trajectories"].shape[1] == cols


def test_curvature_of_circle():
    """Discrete kappa of three points on a circle of radius R (uniform angular steps) equals the
    closed form of the discrete formula; R -> 1/R as the step shrinks."""
    R = 2.0
    om, dt, n = 1.0, 0.01, 50
    N = 64
    c = 32.0
    x = np.arange(N, dtype=np.float64)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    w = np.zeros((3, 3, N, N, N), np.float32)
    w[:, 0] = -om * (Y - c)
    w[:, 1] = om * (X - c)
    w[:, 2] = 0.0
    seeds = np.array([[c + R, c, c]], np.float32)
    st = base3(Integrator=1, DeltaT=dt, MaximumSimulationTime=n * dt, DataTimeDistance=100.0,
               CurvatureMode=3)
    r = oracle.run(st, w, seeds)
    k = r["curvatures"][0]
    assert k.shape == (n - 1,)
    np.testing.assert_allclose(k, 1.0 / R, rtol=1e-4)
    # total curvature ~ kappa * arc length = (1/R) * (R * om * (n - 1) * dt)
    assert r["total_curvature"][0] == pytest.approx(om * (n - 1) * dt, rel=1e-3)


@pytest.mark.parametrize("src,val", [("1-2-3", -4.0), ("2+3*4", 14.0), ("8/4/2", 1.0),
                                     ("(1+2)*3", 9.0), ("2*-3", -6.0), ("10-2*3+1", 5.0),
                                     ("sqrt(16)+ln(1)", 4.0), ("atan2(1, 1)*4", math.pi),
                                     (" 1.5 * 2 ", 3.0)])
def test_expression_precedence(src, val):
    f32, f64 = oracle.eval_expr(src, [0.0] * 19)
    assert f64 == pytest.approx(val, rel=1e-12)
    assert f32 == pytest.approx(val, rel=1e-6)


def test_expression_registers():
    ref = [float(i + 1) for i in range(19)]
    assert oracle.eval_expr("x+y*z", ref)[1] == 1 + 2 * 3
    assert oracle.eval_expr("u*v-w", ref)[1] == 17 * 18 - 19
    assert oracle.eval_expr("cv", ref)[1] == 4
    assert oracle.eval_expr("temp + rho", ref, {"temp": 5, "rho": 13})[1] == 5 + 13


@pytest.mark.parametrize("src", ["velocity", "f0+1", "c2", "abs(x)", "1.", ".5", "x y", "",
                                 "sqrt(x", "unknown"])
def test_expression_errors(src):
    with pytest.raises(RuntimeError):
        oracle.eval_expr(src, [0.0] * 19)


def test_double_gyre_fp32_close_to_fp64():
    wind, gs = synthetic.double_gyre(128, 64, 3)
    seeds = synthetic.uniform_seeds(0, 256, [0.05, 0.05], [1.95, 0.95])
    st = dict(Space=1, Integrator=0, DeltaT=0.01, DataTimeDistance=1.0, GridScale=gs,
              Offset=[0, 0], MaximumSimulationTime=2.0, SaveInterval=10)
    a = oracle.run(st, wind, seeds, real64=False)
    b = oracle.run(st, wind, seeds, real64=True)
    assert np.array_equal(a["status"], b["status"])
    d = np.abs(a["trajectories"] - b["trajectories"]).max()
    assert d < 1e-4


def test_stretched_z_linear_field_closed_form():
    """On a stretched z axis the sampler is linear in physical height inside every layer, so a
    field w = c0 + c1 z is reproduced exactly and the FP64 RK4 path of a particle rising through
    many unequal layers follows z(t) = (z0 + c0 / c1) exp(c1 t) - c0 / c1. A wrong layer index,
    fraction or search direction breaks the linearity at layer edges and fails this."""
    nz, ny, nx, S = 24, 4, 4, 2
    levels = synthetic.stretched_levels(nz, 10.0, 1.15)
    c0, c1 = 0.5, 0.02
    w = (c0 + c1 * levels.astype(np.float64)).astype(np.float32)
    wind = np.zeros((S, 3, nz, ny, nx), np.float32)
    wind[:, 2] = w[None, :, None, None]
    dt, steps = 0.05, 400
    st = dict(Space=0, Integrator=1, DeltaT=dt, DataTimeDistance=1e6, GridScale=[1.0, 1.0, 1.0],
              Offset=[0.0, 0.0, 0.0], MaximumSimulationTime=(steps - 0.5) * dt, SaveInterval=50)
    z0 = np.array([0.3, 2.0, 17.5, 40.0], np.float64)
    seeds = np.stack([np.full(4, 1.5), np.full(4, 1.5), z0], axis=1).astype(np.float32)
    r = oracle.run(st, wind, seeds, real64=True, zlevels=levels)
    assert (r["status"] == 0).all()
    # column k is saved after step 50 k + 1 (the save precedes the time increment)
    t = dt * (1 + 50 * np.arange(r["trajectories"].shape[1] - 1))
    zs = np.asarray(seeds[:, 2], np.float64)
    want = (zs[:, None] + c0 / c1) * np.exp(c1 * t[None, :]) - c0 / c1
    got = r["trajectories"][:, 1:, 2]
    # crossed several layers of different thickness (10 m, 11.5 m, 13.2 m, ...)
    assert got.max() > levels[4] and got.min() < levels[1]
    # the field holds float32 values of the line: the relative error stays at FP32 level
    np.testing.assert_allclose(got, want, rtol=2e-6)
