"""The oracle's feature paths pinned by closed forms and brute force.

Each test derives the expected result from the mathematics of the reference's rule (the cited
lines) and not from the oracle's code: dyadic inputs make the expected values exact in FP64, so
most comparisons are exact or at a few ulps, and each is built so that a plausible restatement
mistake (a wrong sign, a dropped term, an off-by-one step, a wrong index or operand) fails it.
"""
import math

import numpy as np
import pytest

import oracle


def uniform2(nx, ny, S, u, v):
    w = np.zeros((S, 2, ny, nx), np.float32)
    w[:, 0], w[:, 1] = u, v
    return w


def uniform3(n, S, u, v, w_):
    w = np.zeros((S, 3, n, n, n), np.float32)
    w[:, 0], w[:, 1], w[:, 2] = u, v, w_
    return w


def st2(**kw):
    d = dict(Space=1, Integrator=0, DeltaT=0.25, DataTimeDistance=100.0, GridScale=[1.0, 1.0],
             Offset=[0.0, 0.0], MaximumSimulationTime=4.0, SaveInterval=1)
    d.update(kw)
    return d


def st3(**kw):
    d = dict(Space=0, Integrator=0, DeltaT=0.25, DataTimeDistance=100.0, GridScale=[1.0, 1.0, 1.0],
             Offset=[0.0, 0.0, 0.0], MaximumSimulationTime=4.0, SaveInterval=1)
    d.update(kw)
    return d


# --------------------------------------------------------------------------------------------
# 2D loop borders (tpp:20-96): positions wrap modulo the domain, Domain = grid * GridScale
# --------------------------------------------------------------------------------------------
@pytest.mark.parametrize("border,axis", [("LoopX", 0), ("LoopY", 1), ("LoopXY", 2)])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_loop_border_wraps_uniform_drift(border, axis, sign):
    """Uniform drift on an 8 x 4 grid with an offset: every saved position is the start plus
    k * u * dt taken modulo the domain extent (8, 4) on the looped axes, exactly (dyadic
    values). Wrapping at (n - 1) * h, forgetting the offset, or a wrong wrap of negative values
    breaks the equality."""
    nx, ny = 8, 4
    ox, oy = 0.5, -2.0
    u = sign * 1.25 if axis in (0, 2) else 0.0
    v = sign * 0.75 if axis in (1, 2) else 0.0
    w = uniform2(nx, ny, 2, u, v)
    seeds = np.array([[ox + 6.5, oy + 2.0], [ox + 0.25, oy + 3.5]], np.float32)
    steps = 48
    st = st2(BorderMode=border, Offset=[ox, oy], MaximumSimulationTime=steps * 0.25)
    r = oracle.run(st, w, seeds)
    T = r["trajectories"]
    assert T.shape[1] == steps + 1
    assert (r["status"] == 0).all()
    k = np.arange(steps + 1)[None, :]
    x = (seeds[:, 0:1] - ox) + k * u * 0.25
    y = (seeds[:, 1:2] - oy) + k * v * 0.25
    if axis in (0, 2):
        x = np.mod(x, nx)
    if axis in (1, 2):
        y = np.mod(y, ny)
    np.testing.assert_array_equal(T[:, :, 0], x + ox)
    np.testing.assert_array_equal(T[:, :, 1], y + oy)


def test_loop_x_leaves_through_y_as_domain():
    """LoopX wraps x only: a particle drifting out through y = Domain.Y stops with Domain at
    the step its new y exceeds the extent (inclusive edge), frozen at its last position."""
    w = uniform2(8, 4, 2, 1.0, 0.5)
    seeds = np.array([[1.0, 3.0]], np.float32)
    r = oracle.run(st2(BorderMode="LoopX"), w, seeds)
    # y = 3 + 0.125 (k + 1): 4.0 at k = 7 (still inside), 4.125 at k = 8 -> Domain at step 8
    assert r["status"][0] == 2 and r["abort_step"][0] == 8
    np.testing.assert_array_equal(r["trajectories"][0, -1], [1.0 + 8 * 0.25, 4.0])


# --------------------------------------------------------------------------------------------
# sphere coordinates (tpp:98-104) and the pole reflection of LoopXY (tpp:63-96)
# --------------------------------------------------------------------------------------------
def test_sphere_coordinates_closed_form():
    """Uniform wind (u, v) in m/s on a lon/lat grid: dx/dt = u * 180 / (R cos(lat) pi) deg/s and
    dy/dt = v * 180 / (pi R). With v = 0 the latitude is constant and the longitude advances
    linearly; with u = 0 the latitude advances linearly. Checked in FP64 against the closed
    form (a missing degree conversion, a swapped cos argument or a dropped pi fails)."""
    R = 6371000.0
    w = uniform2(36, 18, 2, 20.0, 0.0)
    lat0 = np.array([0.0, 30.0, -55.0, 60.0])
    seeds = np.stack([np.full(4, 100.0), lat0], axis=1).astype(np.float32)
    dt, n = 600.0, 40
    st = st2(GridScale=[10.0, 10.0], Offset=[0.0, -90.0], DeltaT=dt, MaximumSimulationTime=n * dt,
             SphereRadius=R, DataTimeDistance=1e9)
    r = oracle.run(st, w, seeds)
    lat = seeds[:, 1].astype(np.float64)
    rate = 20.0 * 180.0 / (R * np.cos(lat * 2 * math.pi / 360.0) * math.pi)
    k = np.arange(n + 1)[None, :]
    np.testing.assert_allclose(r["trajectories"][:, :, 0], seeds[:, :1] + k * dt * rate[:, None],
                               rtol=1e-12)
    np.testing.assert_array_equal(r["trajectories"][:, :, 1], np.repeat(seeds[:, 1:2], n + 1, axis=1))
    w2 = uniform2(36, 18, 2, 0.0, 15.0)
    r2 = oracle.run(st, w2, seeds)
    vrate = 15.0 * 180.0 / (math.pi * R)
    np.testing.assert_allclose(r2["trajectories"][:, :, 1], seeds[:, 1:2] + k * dt * vrate, rtol=1e-12)


def test_sphere_pole_crossing_reflects_and_turns_longitude():
    """LoopXY on a sphere: a step that crosses the south edge (y < Offset.Y) reflects the
    latitude about the pole and moves the longitude by 180 degrees plus twice the crossing
    longitude's offset (tpp:63-96). Straight southward motion (u = 0) crosses at its own
    longitude, so the new position is (x + 180 mod 360, -(y - oy) + oy)."""
    R = 1000.0
    v = -(R * math.pi / 180.0)  # exactly one degree of latitude per second
    w = uniform2(36, 18, 2, 0.0, v)
    seeds = np.array([[40.0, -89.25]], np.float32)
    st = st2(GridScale=[10.0, 10.0], Offset=[0.0, -90.0], DeltaT=1.0, MaximumSimulationTime=1.0,
             SphereRadius=R, DataTimeDistance=1e9, BorderMode="LoopXY")
    r = oracle.run(st, w, seeds)
    dy = float(np.float32(v)) * 180.0 / (math.pi * R)  # about -1 degree (v is stored as FP32)
    ny_rel = (-89.25 + 90.0) + dy  # -0.25 below the pole
    want = [(40.0 + 180.0) % 360.0, -ny_rel - 90.0]
    np.testing.assert_allclose(r["trajectories"][0, 1], want, rtol=0, atol=1e-12)
    assert r["status"][0] == 0


# --------------------------------------------------------------------------------------------
# Length / FitLength aborts (tpp:253-280)
# --------------------------------------------------------------------------------------------
def test_length_abort_step_and_fit_length_exact():
    """Uniform 3D wind (0.75, 1, 0), |u| dt = 0.3125 per step (exact): Length aborts at counter
    floor(L / 0.3125), the first step whose accumulated length exceeds L, and keeps that step's
    position; FitLength shortens that step so that the length is exactly MaximumLength. Summing
    components instead of the Euclidean length, or testing before adding, moves the step."""
    w = uniform3(8, 2, 0.75, 1.0, 0.0)
    seeds = np.array([[1.0, 1.0, 4.0]], np.float32)
    L = float(np.float32(2.1))
    d = 0.3125
    r = oracle.run(st3(AbortMode="Length", MaximumLength=L), w, seeds)
    k = math.floor(L / d)  # 6: after 7 steps the length is 2.1875 > 2.1
    assert r["status"][0] == 1 and r["abort_step"][0] == k
    assert r["lengths"][0] == (k + 1) * d
    np.testing.assert_array_equal(r["trajectories"][0, -1], [1.0 + (k + 1) * 0.1875, 1.0 + (k + 1) * 0.25, 4.0])
    rf = oracle.run(st3(AbortMode="FitLength", MaximumLength=L), w, seeds)
    assert rf["status"][0] == 1 and rf["abort_step"][0] == k
    assert rf["lengths"][0] == L
    end = rf["trajectories"][0, -1]
    travelled = math.hypot(end[0] - 1.0, end[1] - 1.0)
    assert travelled == pytest.approx(L, rel=1e-14)
    # the last step is shortened along the same direction
    assert (end[1] - 1.0) / (end[0] - 1.0) == pytest.approx(1.0 / 0.75, rel=1e-14)


# --------------------------------------------------------------------------------------------
# uplift (tpp:233-247): added to dp after the velocity step, 3D only
# --------------------------------------------------------------------------------------------
def test_uplift_constant_and_dynamic_closed_form():
    """Zero wind plus per-particle uplift: Constant mode moves z by exactly lift per step;
    Dynamic mode lerps the uplift of the window's two slices in time, so with slices linear
    in time (lift_s = a + b s) the step taken at time t adds a + b t."""
    S, n = 3, 2
    w = uniform3(8, S, 0.0, 0.0, 0.0)
    seeds = np.array([[2.0, 2.0, 1.0], [3.0, 5.0, 2.0]], np.float32)
    lift = np.zeros((S, n, 3), np.float32)
    lift[:, 0, 2] = 0.25
    lift[:, 1, 2] = 0.125
    r = oracle.run(st3(UpLiftMode="Constant", MaximumSimulationTime=2.0), w, seeds, uplift=lift)
    k = np.arange(9)
    np.testing.assert_array_equal(r["trajectories"][0, :, 2], 1.0 + 0.25 * k)
    np.testing.assert_array_equal(r["trajectories"][1, :, 2], 2.0 + 0.125 * k)
    a, b = 0.25, 0.5
    lift_d = np.zeros((S, n, 3), np.float32)
    for s in range(S):
        lift_d[s, :, 2] = a + b * s
    r = oracle.run(st3(UpLiftMode="Dynamic", DataTimeDistance=1.0, MaximumSimulationTime=2.0), w,
                   seeds, uplift=lift_d)
    t = 0.25 * np.arange(8)
    want = seeds[0, 2] + np.concatenate([[0.0], np.cumsum(a + b * t)])
    np.testing.assert_array_equal(r["trajectories"][0, :, 2], want)


# --------------------------------------------------------------------------------------------
# topography (icc:131-143): height bilinear at (x / sx, y / sy), Offset not applied
# --------------------------------------------------------------------------------------------
def test_topography_flat_downdraft_abort_step():
    """Flat ground at height h under a uniform downdraft: the particle stops with Topography
    at the first step whose new z is below h, frozen at the previous position."""
    w = uniform3(8, 2, 0.0, 0.0, -1.0)
    seeds = np.array([[3.0, 3.0, 5.0]], np.float32)
    topo = np.full((8, 8), 2.1, np.float32)
    r = oracle.run(st3(MaximumSimulationTime=10.0), w, seeds, topography=topo)
    # z = 5 - 0.25 (k + 1) < 2.1 first at k + 1 = 12
    assert r["status"][0] == 3 and r["abort_step"][0] == 11
    assert r["trajectories"][0, -1, 2] == 5.0 - 0.25 * 11


def test_topography_ignores_offset():
    """A ground ramp rising with the x index: the height under a particle is read at index
    x / sx (world coordinate, Offset ignored), so with Offset.X = 2 the ramp is met where the
    index equals the world x. Reading at (x - Offset) / sx instead moves the abort by steps."""
    nx = 8
    w = uniform3(nx, 2, 0.5, 0.0, 0.0)
    topo = np.tile(np.arange(nx, dtype=np.float32) * 0.5, (nx, 1))  # height = 0.5 * index
    seeds = np.array([[2.25, 3.0, 2.2]], np.float32)
    st = st3(Offset=[2.0, 0.0, 0.0], MaximumSimulationTime=8.0)
    r = oracle.run(st, w, seeds, topography=topo)
    # new x = 2.25 + 0.125 (k + 1); ground 0.5 x > 2.2 first at x > 4.4 -> k + 1 = 18
    assert r["status"][0] == 3 and r["abort_step"][0] == 17


# --------------------------------------------------------------------------------------------
# reverse (tpp:249, 282): dp = -v dt, new = p - dp -- the motion stays forward
# --------------------------------------------------------------------------------------------
@pytest.mark.parametrize("integ", [0, 1])
def test_reverse_double_negation_moves_forward(integ):
    """Reverse negates dp and subtracts it, so trajectories equal the forward run bit for bit
    (negation is exact) on any field; with uniform wind they move along +u."""
    rng = np.random.default_rng(3)
    w = rng.uniform(-1, 1, (3, 3, 8, 8, 8)).astype(np.float32)
    seeds = rng.uniform(2, 6, (16, 3)).astype(np.float32)
    st = st3(Integrator=integ, DeltaT=0.05, DataTimeDistance=1.0, MaximumSimulationTime=1.5,
             CurvatureMode=3)
    for real64 in (False, True):
        a = oracle.run(st, w, seeds, real64=real64)
        b = oracle.run(dict(st, Reverse=True), w, seeds, real64=real64)
        np.testing.assert_array_equal(a["trajectories"], b["trajectories"])
        np.testing.assert_array_equal(a["status"], b["status"])
    u = uniform3(8, 2, 0.5, 0.0, 0.0)
    r = oracle.run(dict(st, Reverse=True), u, np.array([[2.0, 2.0, 2.0]], np.float32))
    assert r["trajectories"][0, -1, 0] > 2.0


# --------------------------------------------------------------------------------------------
# FastTotalCurvature (tpp:337-346): sum of acos between consecutive velocity directions
# --------------------------------------------------------------------------------------------
def test_fast_total_curvature_full_loop():
    """Solid-body rotation (bilinear, sampled exactly) under RK4: each step multiplies the
    position about the centre by the complex factor g = sum_{k<=4} (i w dt)^k / k!, so the
    RK4 velocity direction turns by arg(g) per step. The fast total curvature adds that angle
    for every step after time 0: (n - 1) arg(g), about 2 pi over one revolution."""
    om, dt = 1.0, 0.02
    N, c = 64, 32.0
    x = np.arange(N, dtype=np.float64)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    w = np.zeros((2, 3, N, N, N), np.float32)
    w[:, 0] = -om * (Y - c)
    w[:, 1] = om * (X - c)
    z = 1j * om * float(np.float32(dt))
    g = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    theta = math.atan2(g.imag, g.real)
    n = int(round(2 * math.pi / theta)) + 1
    st = st3(Integrator=1, DeltaT=dt, MaximumSimulationTime=(n - 0.5) * dt, DataTimeDistance=1e6,
             CurvatureMode="FastTotalCurvature", SaveInterval=0)
    r = oracle.run(st, w, np.array([[c + 5.0, c, c]], np.float32))
    assert r["steps"] == n
    assert r["total_curvature"][0] == pytest.approx((n - 1) * theta, rel=1e-9)
    assert r["total_curvature"][0] == pytest.approx(2 * math.pi, abs=theta)


# --------------------------------------------------------------------------------------------
# additional volumes and constants (tpp:363-439), read by expressions and saved
# --------------------------------------------------------------------------------------------
def test_linear_volume_and_constant_values():
    """A volume field linear in x, y, z is reproduced exactly by trilinear sampling at the new
    position, so the saved volume values and an expression reading it equal a + b x + c y +
    d z of the saved positions; a per-particle constant returns its value. Dynamic volumes
    and constants lerp the window's two slices in time (linear in time here)."""
    n, S = 8, 3
    w = uniform3(n, S, 0.25, 0.5, 0.125)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    base = (1.0 + 0.5 * i + 0.25 * j + 0.125 * k).astype(np.float32)
    vol = np.stack([base + 2.0 * s for s in range(S)]).astype(np.float32)  # (S, nz, ny, nx)
    seeds = np.array([[1.0, 1.5, 2.0], [3.25, 2.0, 1.0]], np.float32)
    const = np.stack([np.array([7.0, -3.5], np.float32) + 1.0 * s for s in range(S)])  # (S, N)
    for mode in ("Constant", "Dynamic"):
        st = st3(DataTimeDistance=1.0, MaximumSimulationTime=1.5, AdditionalVolumeMode=mode,
                 ConstantsMode=mode, AdditionalCompute=["temp*2+rho", "rho"])
        r = oracle.run(st, w, seeds, volumes={"temp": vol}, constants={"rho": const})
        P = r["trajectories"][:, 1:, :]  # saved after each step
        t = 0.25 * (np.arange(P.shape[1]))  # time at the start of each step
        lin = 1.0 + 0.5 * P[..., 0] + 0.25 * P[..., 1] + 0.125 * P[..., 2]
        if mode == "Dynamic":
            lin = lin + 2.0 * t[None, :]
            cv = const[0][:, None] + 1.0 * t[None, :]
        else:
            cv = np.repeat(const[0][:, None], P.shape[1], axis=1)
        np.testing.assert_allclose(r["volumes"]["temp"], lin, rtol=0, atol=1e-12)
        np.testing.assert_allclose(r["compute"][1], cv, rtol=0, atol=1e-12)
        np.testing.assert_allclose(r["compute"][0], 2 * lin + cv, rtol=0, atol=1e-12)


# --------------------------------------------------------------------------------------------
# MinimumAliveParticles (tpp:798-811, ToolKernel.tpp:4-25)
# --------------------------------------------------------------------------------------------
def test_minimum_alive_particles_stops_at_hand_counted_step():
    """Four particles leave an 8-node 2D domain at steps 2, 5 and 8 (one never leaves). The
    alive count is checked before every step: with MinimumAliveParticles = 3 the run stops
    before counter 6 (2 alive after step 5), i.e. after 6 steps; with 2, after 9 steps."""
    w = uniform2(8, 8, 2, 1.0, 0.0)
    # x + 0.25 (k + 1) > 8 first at k = 2, 5, 8
    seeds = np.array([[7.3, 1.0], [6.55, 2.0], [5.8, 3.0], [0.5, 4.0]], np.float32)
    st = st2(MaximumSimulationTime=3.0)
    r = oracle.run(st, w, seeds)
    assert list(r["abort_step"][:3]) == [2, 5, 8]
    assert oracle.run(dict(st, MinimumAliveParticles=3), w, seeds)["steps"] == 6
    assert oracle.run(dict(st, MinimumAliveParticles=2), w, seeds)["steps"] == 9
    assert oracle.run(dict(st, MinimumAliveParticles=1), w, seeds)["steps"] == 12


# --------------------------------------------------------------------------------------------
# precision form: Kahan positions, FP64 time
# --------------------------------------------------------------------------------------------
def test_kahan_sum_matches_correctly_rounded_sum():
    """Many tiny equal FP32 increments on a large coordinate: the Kahan-compensated FP32 path
    stays within one ulp of the correctly rounded exact sum, while the plain FP32 sum drifts
    by hundreds of ulps (so the compensation is applied, with the right sign)."""
    w = uniform3(8, 2, 1.0, 0.0, 0.0)
    seeds = np.array([[1.0, 2.0, 2.0]], np.float32)
    dt, n = 1.0e-4, 20000
    st = st3(DeltaT=dt, MaximumSimulationTime=(n - 0.5) * dt, SaveInterval=0, DataTimeDistance=1e6)
    inc = np.float64(np.float32(1.0) * np.float32(dt))  # the FP32 increment u * dt
    exact = 1.0 + n * inc
    want = np.float32(exact)
    ulp = np.spacing(want)
    k = oracle.run(st, w, seeds, real64=False, kahan=True)
    p = oracle.run(st, w, seeds, real64=False)
    assert k["steps"] == n
    assert abs(np.float64(k["trajectories"][0, -1, 0]) - exact) <= ulp
    assert abs(np.float64(p["trajectories"][0, -1, 0]) - exact) > 100 * ulp


def _steps_per_window(dt, dtd, nwin, acc):
    """The host loop's window edges (Time += dt, window ends at Time >= max_windfield_time,
    max_windfield_time += DTD) with Time accumulated in `acc` (numpy float32 or float64)."""
    out, t, mwt = [], acc(0), acc(dtd)
    for _ in range(nwin):
        k = 0
        while True:
            k += 1
            t = acc(t + acc(np.float32(dt)))
            if t >= mwt:
                break
        out.append(k)
        mwt = acc(mwt + acc(np.float32(dtd)))
    return out


def test_time64_window_edges_follow_fp64_sequence():
    """With FP64 time the window edges fall where k * float32(0.01) first reaches c + 1 in
    double arithmetic (101, 100, 100, ... steps); FP32 time gives 101, 100, 100, 100, 99."""
    w = uniform3(8, 7, 0.1, 0.0, 0.0)
    seeds = np.array([[1.0, 1.0, 1.0]], np.float32)
    st = st3(DeltaT=0.01, DataTimeDistance=1.0, MaximumSimulationTime=100.0, SaveInterval=0)
    w64 = _steps_per_window(0.01, 1.0, 6, np.float64)
    w32 = _steps_per_window(0.01, 1.0, 6, np.float32)
    assert w64 == [101, 100, 100, 100, 100, 100]
    assert w32[:5] == [101, 100, 100, 100, 99]
    assert oracle.run(st, w, seeds, time64=True)["steps"] == sum(w64)
    assert oracle.run(st, w, seeds)["steps"] == sum(w32)


# --------------------------------------------------------------------------------------------
# texture-weight emulation: 9-bit fixed point weights with 8 fractional bits
# --------------------------------------------------------------------------------------------
@pytest.mark.parametrize("frac,round_w,trunc_w", [(0.3, 77, 76), (0.5, 128, 128), (0.999, 256, 255),
                                                   (0.0019, 0, 0), (0.002, 1, 0)])
def test_q8_weights_brute_force(frac, round_w, trunc_w):
    """A 2D ramp u = i sampled at x = frac returns the weight itself: exact weights give
    float32(frac); q8 rounding gives floor(256 frac + 1/2) / 256 and truncation floor(256 frac)
    / 256 (brute-forced here from the fraction)."""
    nx = 4
    w = np.zeros((2, 2, 4, nx), np.float32)
    w[:, 0] = np.arange(nx, dtype=np.float32)[None, None, :]
    w[:, 1] = 1.0
    seeds = np.array([[frac, 1.0]], np.float32)
    dt = 2.0 ** -10
    st = st2(DeltaT=dt, MaximumSimulationTime=dt)
    f32 = float(np.float32(frac))
    assert math.floor(f32 * 256 + 0.5) == round_w and math.floor(f32 * 256) == trunc_w
    for mode, want in (("exact", f32), ("q8_round", round_w / 256), ("q8_trunc", trunc_w / 256)):
        r = oracle.run(st, w, seeds, weights=mode)
        got = (r["trajectories"][0, 1, 0] - np.float64(np.float32(frac))) / dt
        assert got == pytest.approx(want, abs=1e-6), mode
