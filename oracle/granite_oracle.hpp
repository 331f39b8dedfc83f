// ORACLE — test infrastructure only. The product (pygranite_amd) never links, imports or runs
// anything under oracle/. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// use it, and only as the checker / the CPU baseline.
//
// CPU restatement of pygranite's trajectory integrator (v1.5.0), templated on the arithmetic
// type: R = double is the FP64 oracle, R = float is the plain FP32 CPU path. Every function
// names the reference lines it follows (paths relative to the reference's repository root).
//
//   step body ............... source/granite/TrajectoryIntegrator.tpp:108-457 (integrate)
//   sampling ................ source/granite/TrajectoryIntegrator.icc:22-129
//   domain / topography ..... source/granite/TrajectoryIntegrator.icc:7-20, 131-143
//   curvature ............... source/granite/TrajectoryIntegrator.icc:145-166, tpp:335-361
//   2D loop borders, sphere . source/granite/TrajectoryIntegrator.tpp:13-104
//   schedule / saves ........ source/granite/TrajectoryIntegrator.tpp:778-995, 997-1049
//   setup ................... source/granite/TrajectoryIntegrator.tpp:467-715
//   math helpers ............ source/my/math/LinearAlgebra.hpp:553-654
//
// Texture sampling is restated from the CUDA Programming Guide's description of linear
// filtering with unnormalised coordinates (texel i at world Offset + i*GridScale, weight =
// frac(arg)), as nested lerps x -> y -> z. Weights are exact, or quantised to 1/256
// (round half up, or truncate) to emulate the texture unit's 8-bit fractional weights.
// Addressing: 3D = border (zero outside), 2D = clamp to edge (source/granite/
// TrajectoryIntegrator.cpp:503-506 and 570-574).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "expr_oracle.hpp"

namespace oracle
{

enum WeightMode : int { W_EXACT = 0, W_Q8_ROUND = 1, W_Q8_TRUNC = 2 };
enum Status : int32_t { S_TIME = 0, S_LENGTH = 1, S_DOMAIN = 2, S_TOPO = 3, S_WIND = 4 };

struct Settings
{
    int dim{3};             // 2 or 3
    int integrator{1};      // 0 ExplicitEuler, 1 ClassicRungeKutta
    int border{0};          // 0 Block, 1 LoopX, 2 LoopY, 3 LoopXY (2D only)
    int curvature{0};       // 0 Off, 1 Total, 2 FastTotal, 3 IndividualAndTotal
    int abort_mode{0};      // 0 Time, 1 Length, 2 FitLength
    int uplift_mode{0};     // 0 Off, 1 Constant, 2 Dynamic (3D only)
    int windfield_mode{1};  // 0 Constant, 1 Dynamic
    int constants_mode{0};  // 0 Off, 1 Constant, 2 Dynamic
    int volume_mode{0};     // 0 Off, 1 Constant, 2 Dynamic
    int reverse{0};
    int weight_mode{W_EXACT};
    int time64{0};       // accumulate simulation time in double (precision form)
    int kahan{0};        // Kahan-compensated position update (precision form)
    float sphere_radius{0.f};
    float dt{1.f}, dtd{60.f};
    float max_length{0.f}, max_sim_time{100.f};
    float grid_scale[3]{25.f, 25.f, 25.f};
    float offset[3]{0.f, 0.f, 0.f};
    uint64_t save_interval{1};
    int32_t min_alive{0};
    int nx{0}, ny{0}, nz{1};
    int nslices{0};
    int nexpr{0};
    std::vector<Program> exprs;
    int nconst{0}, nvol{0};
    int have_topo{0};
    // stretched z axis (extension): node k at height zlev[k] relative to Offset.Z
    std::vector<float> zlev;
};

// One integration step of the host schedule.
struct StepInfo
{
    float time32;        // data.Time as the reference accumulates it (FP32)
    double time64;       // same, accumulated in double (used when time64 is set)
    uint32_t window;     // COUNTER of the window (back slice index)
    uint64_t counter;    // _simulation_counter before the step
    int64_t save_col;    // index of the saved column written after this step, or -1
};

// Field data for the whole run: slice s, component c at (k, j, i) lives at
// wind[((s * dim + c) * nz + k) * ny * nx + j * nx + i] (2D: nz = 1).
struct Inputs
{
    const float * wind{nullptr};
    const float * topo{nullptr};      // [ny][nx]
    const float * uplift{nullptr};    // [nslices][N][dim]
    const float * constants{nullptr}; // [nslices][nconst][N]
    const float * volumes{nullptr};   // [nslices][nvol][nz][ny][nx]
};

template <class R> struct Outputs
{
    // positions [col][N][dim] (col 0 = start set), curvature [ccol][N], expressions
    // [col][e][N], volumes [col][v][N] (expression/volume columns exclude the start set)
    std::vector<R> positions;
    std::vector<R> curvature;
    std::vector<R> exprs;
    std::vector<R> volumes;
    std::vector<int32_t> status;
    std::vector<int64_t> abort_step; // counter of the step that changed the status, or -1
    std::vector<R> total_curvature;
    std::vector<R> lengths;
    uint64_t steps{0};
    uint64_t cols{0}, curv_cols{0};
};

inline bool any_dynamic(const Settings & s)
{
    return s.windfield_mode == 1 || s.uplift_mode == 2 || s.constants_mode == 2 ||
           s.volume_mode == 2;
}

// -------------------------------------------------------------------------------------------
// Sampling. arg is in texel units: arg = (p - Offset) * (1 / GridScale)  (icc:112-129).
// -------------------------------------------------------------------------------------------
// a + t * (b - a) (LinearAlgebra.hpp:617-621) evaluated as fma(t, b - a, a): the form nvcc
// compiles it to in the reference's kernels (multiply-add contraction is nvcc's default)
template <class R> inline R lerp(R a, R b, R t) { return std::fma(t, b - a, a); }

template <class R> inline R quant(R w, int mode)
{
    if (mode == W_Q8_ROUND) return std::floor(w * R(256) + R(0.5)) / R(256);
    if (mode == W_Q8_TRUNC) return std::floor(w * R(256)) / R(256);
    return w;
}

// Splits a texel coordinate into (index, weight). The coordinate is first clamped to
// [-1, n]; outside that range every tap is a border/edge texel, so the clamp does not change
// the result, and it keeps the float -> int conversion defined (NaN maps to -1).
template <class R> inline void split(R a, int n, int mode, int & i, R & w)
{
    a = std::fmin(std::fmax(a, R(-1)), R(n));
    R f = std::floor(a);
    i = static_cast<int>(f);
    w = quant(a - f, mode);
}

template <class R> struct Sampler
{
    int nx, ny, nz, dim, mode;

    // 3D border addressing: taps outside [0, n-1] read 0.
    R tap3(const float * g, int i, int j, int k) const
    {
        if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) return R(0);
        return R(g[(static_cast<size_t>(k) * ny + j) * nx + i]);
    }
    // 2D clamp addressing.
    R tap2(const float * g, int i, int j, int w, int h) const
    {
        i = std::min(std::max(i, 0), w - 1);
        j = std::min(std::max(j, 0), h - 1);
        return R(g[static_cast<size_t>(j) * w + i]);
    }

    R trilinear(const float * g, R ax, R ay, R az) const
    {
        int i, j, k;
        R wx, wy, wz;
        split(ax, nx, mode, i, wx);
        split(ay, ny, mode, j, wy);
        split(az, nz, mode, k, wz);
        R c00 = lerp(tap3(g, i, j, k), tap3(g, i + 1, j, k), wx);
        R c10 = lerp(tap3(g, i, j + 1, k), tap3(g, i + 1, j + 1, k), wx);
        R c01 = lerp(tap3(g, i, j, k + 1), tap3(g, i + 1, j, k + 1), wx);
        R c11 = lerp(tap3(g, i, j + 1, k + 1), tap3(g, i + 1, j + 1, k + 1), wx);
        R c0 = lerp(c00, c10, wy);
        R c1 = lerp(c01, c11, wy);
        return lerp(c0, c1, wz);
    }

    R bilinear(const float * g, R ax, R ay, int w, int h) const
    {
        int i, j;
        R wx, wy;
        split(ax, w, mode, i, wx);
        split(ay, h, mode, j, wy);
        R c0 = lerp(tap2(g, i, j, w, h), tap2(g, i + 1, j, w, h), wx);
        R c1 = lerp(tap2(g, i, j + 1, w, h), tap2(g, i + 1, j + 1, w, h), wx);
        return lerp(c0, c1, wy);
    }

    R sample(const float * g, const R * arg) const
    {
        return dim == 3 ? trilinear(g, arg[0], arg[1], arg[2]) : bilinear(g, arg[0], arg[1], nx, ny);
    }
};

// -------------------------------------------------------------------------------------------
// Host schedule: tpp:778-995 (computeLoop_ / computeLoopGPU_). Walks the windows and steps in
// order, calling `alive(counter)` before every step when MinimumAliveParticles is set and
// `on_step(info)` for every step. Returns the FP32 simulation time after the last step.
// -------------------------------------------------------------------------------------------
template <class Alive, class OnStep>
inline float walk_schedule(const Settings & s, Alive && alive, OnStep && on_step)
{
    const bool dyn = any_dynamic(s);
    const float max_steps_f = s.max_sim_time / s.dt; // size_t < float compare (tpp:974)
    uint64_t counter = 0;
    float time32 = 0.f;
    double time64 = 0.0;
    float max_wt = s.dtd;      // max_windfield_time (tpp:783), += DTD per window (tpp:892)
    double max_wt64 = s.dtd;
    uint32_t c = 0;            // window COUNTER (tpp:902)
    int64_t col = 0;
    bool abort_due_alive = false;
    for (;;)
    {
        // Dynamic modes run the window on the calling thread with max_windfield_time; the
        // all-constant case runs one "window" bounded by MaximumSimulationTime (tpp:857-880).
        const float mwt = dyn ? max_wt : s.max_sim_time;
        const double mwt64 = dyn ? max_wt64 : double(s.max_sim_time);
        while (static_cast<float>(counter) < max_steps_f)
        {
            if (s.min_alive != 0 && !alive(counter))
            {
                abort_due_alive = true;
                break;
            }
            StepInfo si{time32, time64, c, counter, -1};
            if (s.save_interval != 0 && counter % s.save_interval == 0) si.save_col = col++;
            on_step(si);
            time32 += s.dt;
            time64 += double(s.dt);
            counter++;
            if (s.time64 ? (time64 >= mwt64) : (time32 >= mwt)) break;
        }
        if (dyn)
        {
            // _loader.step() succeeds while slice c + 2 exists (tpp:887-908)
            if (static_cast<int>(c) + 2 < s.nslices)
            {
                if (abort_due_alive) break;
                max_wt += s.dtd;
                max_wt64 += double(s.dtd);
                c++;
            }
            else
                break;
        }
        if (!(s.windfield_mode == 1 && static_cast<float>(counter) < max_steps_f)) break;
    }
    return time32;
}

inline std::vector<StepInfo> schedule(const Settings & s)
{
    std::vector<StepInfo> steps;
    walk_schedule(s, [](uint64_t) { return true; }, [&](const StepInfo & si) { steps.push_back(si); });
    return steps;
}

// -------------------------------------------------------------------------------------------
// Integrator
// -------------------------------------------------------------------------------------------
template <class R> class Integrator
{
    static constexpr bool kFloat = sizeof(R) == sizeof(float);

public:
    Integrator(const Settings & s, const Inputs & in, const float * seeds, size_t n)
        : s_(s), in_(in), n_(n)
    {
        D_ = s.dim;
        sp_ = Sampler<R>{s.nx, s.ny, s.dim == 3 ? s.nz : 1, s.dim, s.weight_mode};
        // GridScale/Offset arrive already expanded to X, Y, Z (tpp:521-574).
        for (int d = 0; d < 3; ++d)
        {
            h_[d] = R(s.grid_scale[d]);
            inv_h_[d] = R(1) / h_[d];
            off_[d] = R(s.offset[d]);
        }
        // Domain = grid * GridScale computed in float (tpp:526-546); on a stretched z axis
        // (extension) the z extent is the top level.
        dom_[0] = R(float(s.nx) * s.grid_scale[0]);
        dom_[1] = R(float(s.ny) * s.grid_scale[1]);
        dom_[2] = s.zlev.empty() ? R(float(s.nz) * s.grid_scale[2]) : R(s.zlev.back());
        pos_.assign(n * 3, R(0));
        prev_.assign(n * 3, R(0));
        comp_.assign(n * 3, R(0));
        for (size_t i = 0; i < n; ++i)
            for (int d = 0; d < D_; ++d) pos_[i * 3 + d] = R(seeds[i * D_ + d]);
        status_.assign(n, S_TIME);
        abort_step_.assign(n, -1);
        total_.assign(n, R(0));
        curv_.assign(n, R(0));
        len_.assign(n, R(0));
        lastdir_.assign(n * 3, R(0));
        expr_.assign(size_t(s.nexpr) * n, R(0));
        vol_.assign(size_t(s.nvol) * n, R(0));
        slab_ = size_t(s.nx) * s.ny * (s.dim == 3 ? s.nz : 1);
    }

    Outputs<R> run(int nthreads)
    {
        Outputs<R> out;
        // initial row (the start set is the first trajectory point, cpp:424-442)
        for (size_t i = 0; i < n_; ++i)
            for (int d = 0; d < D_; ++d) out.positions.push_back(pos_[i * 3 + d]);
        out.cols = 1;
        auto alive = [&](uint64_t) {
            int32_t cnt = 0;
            for (size_t i = 0; i < n_; ++i) cnt += status_[i] == S_TIME;
            return cnt >= s_.min_alive; // count_if + compare (tpp:803-811)
        };
        auto on_step = [&](const StepInfo & st) {
            par(nthreads, [&](size_t b, size_t e) {
                for (size_t i = b; i < e; ++i) step(i, st);
            });
            out.steps++;
            if (st.save_col >= 0) save(out, st.time32 > 0.f);
        };
        const float t_end = walk_schedule(s_, alive, on_step);
        if (s_.save_interval == 0) save(out, t_end > 0.f); // tpp:914-915

        out.status = status_;
        out.abort_step = abort_step_;
        if (s_.curvature != 0) out.total_curvature = total_;
        if (s_.abort_mode != 0) out.lengths = len_;
        return out;
    }

private:
    const Settings & s_;
    Inputs in_;
    size_t n_;
    int D_;
    Sampler<R> sp_;
    R h_[3], inv_h_[3], off_[3], dom_[3];
    std::vector<R> pos_, prev_, comp_;
    std::vector<int32_t> status_;
    std::vector<int64_t> abort_step_;
    std::vector<R> total_, curv_, len_, lastdir_, expr_, vol_;
    size_t slab_{0};

    template <class F> static void par(int nthreads, F && f, size_t n)
    {
        if (nthreads <= 1 || n < 4096)
        {
            f(size_t(0), n);
            return;
        }
        std::vector<std::thread> th;
        size_t chunk = (n + nthreads - 1) / nthreads;
        for (int t = 0; t < nthreads; ++t)
        {
            size_t b = t * chunk, e = std::min(n, b + chunk);
            if (b >= e) break;
            th.emplace_back([&f, b, e] { f(b, e); });
        }
        for (auto & x : th) x.join();
    }
    template <class F> void par(int nthreads, F && f) { par(nthreads, std::forward<F>(f), n_); }

    const float * wind(uint32_t slice, int comp) const
    {
        // Constant windfield mode always samples the first slice (tpp:816-821)
        if (s_.windfield_mode == 0) slice = 0;
        return in_.wind + (size_t(slice) * D_ + comp) * slab_;
    }

    // texture_arg (icc:112-129); on a stretched z axis the height maps to
    // k + (z - z_k) / (z_{k+1} - z_k) with z_k <= z < z_{k+1} (first / last cell continued
    // linearly outside the levels) -- found here by binary search
    void texel_arg(const R * p, R * arg) const
    {
        for (int d = 0; d < D_; ++d) arg[d] = (p[d] - off_[d]) * inv_h_[d];
        if (D_ == 3 && !s_.zlev.empty())
        {
            const std::vector<float> & z = s_.zlev;
            const int nz = int(z.size());
            const R zr = p[2] - off_[2];
            int k = 0;
            if (zr >= R(z[0]))
            {
                // last level <= zr, at most nz - 2
                int lo = 0, hi = nz - 2;
                while (lo < hi)
                {
                    const int mid = (lo + hi + 1) / 2;
                    if (R(z[size_t(mid)]) <= zr)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                k = lo;
            }
            const R z0 = R(z[size_t(k)]), z1 = R(z[size_t(k) + 1]);
            arg[2] = R(k) + (zr - z0) / (z1 - z0);
        }
    }

    // interpolate(data, p, t) icc:42-50 / 70-77; constant mode icc:79-87 / 104-110
    void velocity(const R * p, R t, uint32_t back, uint32_t front, bool timed, R * v) const
    {
        R arg[3] = {R(0), R(0), R(0)};
        texel_arg(p, arg);
        for (int c = 0; c < D_; ++c)
        {
            if (timed)
                v[c] = lerp(sp_.sample(wind(back, c), arg), sp_.sample(wind(front, c), arg), t);
            else
                v[c] = sp_.sample(wind(back, c), arg);
        }
        if (D_ == 2) v[2] = R(0);
    }

    // sphere_coordinates tpp:98-104
    void sphere(R py, const R * v, R * o) const
    {
        // the reference's float constants (my/math/Constants.hpp:9-14); exact pi in FP64
        const R PI = kFloat ? R(3.14159265358979323846f) : R(3.14159265358979323846);
        const R PI_2 = kFloat ? R(3.14159265358979323846f * 2.f) : R(2) * PI;
        const R rad = R(s_.sphere_radius);
        o[0] = (v[0] * R(180)) / (rad * std::cos(py * PI_2 / R(360)) * PI);
        o[1] = (v[1] * R(180)) / (PI * rad);
        o[2] = R(0);
    }

    // unit_interval(Time, TimeFront, TimeBack) = (Time - TimeBack) / (TimeFront - TimeBack)
    R unit(const StepInfo & st, R offset) const
    {
        const float tb = float(st.window) * s_.dtd, tf = float(st.window + 1) * s_.dtd;
        if (s_.time64)
        {
            const double tb64 = double(st.window) * double(s_.dtd);
            const double tf64 = double(st.window + 1) * double(s_.dtd);
            double x = (st.time64 + double(offset) - tb64) / (tf64 - tb64);
            return R(static_cast<float>(x));
        }
        return (R(st.time32) + offset - R(tb)) / (R(tf) - R(tb));
    }

    bool in_topography(const R * p) const
    {
        if (!in_.topo) return false;
        const R ax = p[0] / h_[0], ay = p[1] / h_[1];
        return sp_.bilinear(in_.topo, ax, ay, s_.nx, s_.ny) > p[2];
    }

    bool out_of_box(const R * p) const
    {
        for (int d = 0; d < D_; ++d)
            if (p[d] - off_[d] > dom_[d] || p[d] - off_[d] < R(0)) return true;
        return false;
    }

    static R fmod_(R a, R b) { return std::fmod(a, b); }

    void step(size_t i, const StepInfo & st)
    {
        R * P = &pos_[i * 3];
        if (status_[i] != S_TIME) return; // dead particles keep their position (tpp:452-455)
        const R dt = R(s_.dt);
        const bool dyn = any_dynamic(s_);
        const uint32_t back = st.window, front = st.window + 1;
        R t = R(0);
        if (dyn) t = unit(st, R(0));
        R v[3] = {R(0), R(0), R(0)}, vc[3] = {R(0), R(0), R(0)};
        const bool timed = s_.windfield_mode == 1;
        const bool sph = s_.sphere_radius > 0.f && D_ == 2;
        if (s_.integrator == 1)
        {
            R t1 = t, t2 = t;
            if (timed)
            {
                t1 = unit(st, R(0.5f) * dt);
                t2 = unit(st, dt);
            }
            R k1[3], k2[3], k3[3], k4[3], q[3] = {P[0], P[1], P[2]};
            R k1c[3], k2c[3], k3c[3], k4c[3];
            auto f = [&](const R * x, R tt, R * kc, R * k) {
                velocity(x, tt, back, front, timed, kc);
                if (sph)
                    sphere(x[1], kc, k);
                else
                    for (int d = 0; d < 3; ++d) k[d] = kc[d];
            };
            f(P, t, k1c, k1);
            for (int d = 0; d < D_; ++d) q[d] = P[d] + k1[d] * (R(0.5f) * dt);
            f(q, t1, k2c, k2);
            for (int d = 0; d < D_; ++d) q[d] = P[d] + k2[d] * (R(0.5f) * dt);
            f(q, t1, k3c, k3);
            for (int d = 0; d < D_; ++d) q[d] = P[d] + k3[d] * dt;
            f(q, t2, k4c, k4);
            // (1.0f / 6) in FP32 as in tpp:175; the exact 1/6 in FP64
            const R sixth = kFloat ? R(1.0f / 6) : R(1) / R(6);
            for (int d = 0; d < D_; ++d)
            {
                v[d] = (((k1[d] + k2[d] * R(2)) + k3[d] * R(2)) + k4[d]) * sixth;
                vc[d] = sph ? (((k1c[d] + k2c[d] * R(2)) + k3c[d] * R(2)) + k4c[d]) * sixth : v[d];
            }
        }
        else
        {
            velocity(P, t, back, front, timed, vc);
            for (int d = 0; d < 3; ++d) v[d] = vc[d];
            if (sph) sphere(P[1], vc, v);
        }

        // dp (tpp:233-250); uplift is 3D only (TrajectoryIntegratorMacros.hpp:67)
        R dp[3] = {R(0), R(0), R(0)};
        const R sgn = s_.reverse ? R(-1) : R(1);
        for (int d = 0; d < D_; ++d) dp[d] = v[d] * dt * sgn;
        if (s_.uplift_mode != 0 && D_ == 3 && in_.uplift)
        {
            const size_t nd = size_t(D_);
            const float * lb = in_.uplift + (size_t(s_.uplift_mode == 1 ? 0 : back) * n_ + i) * nd;
            const float * lf = in_.uplift + (size_t(front) * n_ + i) * nd;
            for (int d = 0; d < D_; ++d)
                dp[d] = dp[d] + (s_.uplift_mode == 1 ? R(lb[d]) : lerp(R(lb[d]), R(lf[d]), t));
        }

        int32_t abort = S_TIME;
        if (s_.abort_mode != 0) // tpp:253-280
        {
            R l2 = R(0);
            for (int d = 0; d < D_; ++d) l2 += dp[d] * dp[d];
            const R dp_l = std::sqrt(l2);
            const R c_l = len_[i];
            const R n_l = c_l + dp_l;
            const R maxl = R(s_.max_length);
            if (s_.abort_mode == 2)
            {
                if (n_l > maxl)
                {
                    const R sc = (maxl - c_l) / dp_l;
                    for (int d = 0; d < D_; ++d) dp[d] = sc * dp[d];
                    len_[i] = maxl;
                    abort = S_LENGTH;
                }
                else
                    len_[i] = n_l;
            }
            else
            {
                if (n_l > maxl) abort = S_LENGTH;
                len_[i] = n_l;
            }
        }

        R np[3] = {P[0], P[1], P[2]};
        if (s_.kahan)
        {
            // Kahan-compensated position sum: comp_ holds the running lost low-order part.
            for (int d = 0; d < D_; ++d)
            {
                R y = (s_.reverse ? -dp[d] : dp[d]) - comp_[i * 3 + d];
                R tt = P[d] + y;
                np[d] = tt;
            }
        }
        else
            for (int d = 0; d < D_; ++d) np[d] = s_.reverse ? P[d] - dp[d] : P[d] + dp[d];

        R dp2 = R(0);
        for (int d = 0; d < D_; ++d) dp2 += dp[d] * dp[d];
        const R ABORT_VALUE = R(1e-9f);
        if (abort == S_TIME) // tpp:284-323
        {
            if (D_ == 2 && s_.border != 0)
            {
                if (dp2 < ABORT_VALUE)
                    abort = S_WIND;
                else if (loop_border(P, np))
                    abort = S_DOMAIN;
            }
            else
            {
                if (D_ == 3 && s_.have_topo && in_topography(np))
                    abort = S_TOPO;
                else if (dp2 < ABORT_VALUE)
                    abort = S_WIND;
                else if (out_of_box(np))
                    abort = S_DOMAIN;
            }
        }
        status_[i] = abort;
        if (abort != S_TIME) abort_step_[i] = int64_t(st.counter);

        const R time_now = s_.time64 ? R(st.time64) : R(st.time32);
        if (abort == S_WIND || abort == S_TOPO || abort == S_DOMAIN)
        {
            // front = particle; TotalCurvature = -1 (tpp:327-332)
            if (s_.curvature != 0) total_[i] = R(-1);
            for (int d = 0; d < 3; ++d) prev_[i * 3 + d] = P[d];
            return;
        }

        if (s_.curvature != 0)
        {
            if (s_.curvature == 2) // FastTotalCurvature (tpp:337-346)
            {
                R l2 = R(0);
                for (int d = 0; d < D_; ++d) l2 += v[d] * v[d];
                const R inv = R(1) / std::sqrt(l2);
                R vn[3] = {R(0), R(0), R(0)};
                for (int d = 0; d < D_; ++d) vn[d] = v[d] * inv;
                if (time_now > R(0))
                {
                    R dd = R(0);
                    for (int d = 0; d < D_; ++d) dd += lastdir_[i * 3 + d] * vn[d];
                    const R k = std::acos(dd);
                    if (!std::isnan(k)) total_[i] += k;
                }
                for (int d = 0; d < 3; ++d) lastdir_[i * 3 + d] = vn[d];
            }
            else if (time_now > R(0)) // tpp:347-360
            {
                const R * pm1 = &prev_[i * 3];
                R k = R(0);
                if (D_ == 3) k = std::fabs(kappa(pm1, P, np, dt));
                R a2 = R(0), b2 = R(0);
                for (int d = 0; d < D_; ++d)
                {
                    a2 += (P[d] - pm1[d]) * (P[d] - pm1[d]);
                    b2 += (np[d] - P[d]) * (np[d] - P[d]);
                }
                const R ds = R(0.5f) * (std::sqrt(a2) + std::sqrt(b2));
                if (s_.curvature == 3) curv_[i] = k;
                if (!std::isnan(k)) total_[i] += k * ds;
            }
        }

        if (s_.volume_mode != 0 && in_.volumes) // tpp:363-377 (volume sampled at new_pos)
        {
            R arg[3] = {R(0), R(0), R(0)};
            texel_arg(np, arg);
            for (int k = 0; k < s_.nvol; ++k)
            {
                const float * vb =
                    in_.volumes + (size_t(s_.volume_mode == 1 ? 0 : back) * s_.nvol + k) * slab_;
                R val = sp_.sample(vb, arg);
                if (s_.volume_mode == 2)
                {
                    const float * vf = in_.volumes + (size_t(front) * s_.nvol + k) * slab_;
                    val = lerp(val, sp_.sample(vf, arg), t);
                }
                vol_[size_t(k) * n_ + i] = val;
            }
        }

        if (s_.nexpr > 0) // tpp:379-447
        {
            R ref[19];
            for (int r = 0; r < 19; ++r) ref[r] = R(0);
            ref[0] = np[0];
            ref[1] = np[1];
            ref[2] = D_ == 3 ? np[2] : R(0);
            ref[3] = s_.curvature == 3 ? curv_[i] : R(0);
            for (int k = 0; k < s_.nvol && k < 8; ++k) ref[4 + k] = vol_[size_t(k) * n_ + i];
            for (int k = 0; k < s_.nconst && k < 4; ++k)
            {
                const float * cb =
                    in_.constants + (size_t(s_.constants_mode == 2 ? back : 0) * s_.nconst + k) * n_;
                R val = R(cb[i]);
                if (s_.constants_mode == 2)
                {
                    const float * cf = in_.constants + (size_t(front) * s_.nconst + k) * n_;
                    val = lerp(R(cb[i]), R(cf[i]), t);
                }
                ref[12 + k] = val;
            }
            ref[16] = vc[0];
            ref[17] = vc[1];
            ref[18] = D_ == 3 ? vc[2] : R(0);
            for (int e = 0; e < s_.nexpr; ++e)
                expr_[size_t(e) * n_ + i] = eval<R>(s_.exprs[size_t(e)], s_.exprs[size_t(e)].root, ref);
        }

        // commit: prev <- p, p <- new_pos
        if (s_.kahan)
        {
            for (int d = 0; d < D_; ++d)
            {
                R y = (s_.reverse ? -dp[d] : dp[d]) - comp_[i * 3 + d];
                comp_[i * 3 + d] = (np[d] - P[d]) - y;
            }
        }
        for (int d = 0; d < 3; ++d)
        {
            prev_[i * 3 + d] = P[d];
            P[d] = np[d];
        }
    }

    // 2D loop borders, tpp:20-96. Returns true on Domain abort; may wrap np in place.
    bool loop_border(const R * o, R * n) const
    {
        const R ox = off_[0], oy = off_[1], DX = dom_[0], DY = dom_[1];
        const bool sph = s_.sphere_radius > 0.f;
        auto oob = [&](const R * p) {
            return p[0] - ox > DX || p[0] - ox < R(0) || p[1] - oy > DY || p[1] - oy < R(0);
        };
        if (s_.border == 1)
        {
            n[0] = fmod_((n[0] - ox + DX), DX) + ox;
            return n[1] - oy > DY || n[1] - oy < R(0);
        }
        if (s_.border == 2)
        {
            if (sph) return oob(n);
            n[1] = fmod_((n[1] - oy + DY), DY) + oy;
            return n[0] - ox > DX || n[0] - ox < R(0);
        }
        // LoopXY
        if (sph)
        {
            const R nx_ = n[0] - ox, ny_ = n[1] - oy;
            if (ny_ < R(0))
            {
                const R ox_ = o[0] - ox, oy_ = o[1] - oy;
                const R tx = (-oy_) / (ny_ - oy_) * (nx_ - ox_) + ox_;
                n[1] = -(ny_) + oy;
                n[0] = fmod_(DX + R(180.0f) - nx_ + R(2.f) * tx, DX) + ox;
            }
            else if (ny_ >= DY)
            {
                const R ox_ = o[0] - ox, oy_ = o[1] - oy;
                const R tx = (DY - oy_) / (ny_ - oy_) * (nx_ - ox_) + ox_;
                n[1] = (ny_ - DY) + oy;
                n[0] = fmod_(DX + R(180.0f) - nx_ + R(2.f) * tx, DX) + ox;
            }
        }
        else
            n[1] = fmod_((n[1] - oy + DY), DY) + oy;
        n[0] = fmod_((n[0] - ox + DX), DX) + ox;
        return false;
    }

    // kappa icc:145-159
    static R kappa(const R * pm1, const R * p, const R * pp1, R DT)
    {
        R d1[3], d2[3];
        for (int d = 0; d < 3; ++d)
        {
            d1[d] = (pp1[d] - pm1[d]) / (R(2) * DT);
            d2[d] = ((pm1[d] + pp1[d]) - R(2) * p[d]) / (DT * DT);
        }
        const R ldp = std::sqrt((d1[0] * d1[0] + d1[1] * d1[1]) + d1[2] * d1[2]);
        const R ldp3 = (ldp * ldp) * ldp;
        const R cx = d1[1] * d2[2] - d1[2] * d2[1];
        const R cy = d1[2] * d2[0] - d1[0] * d2[2];
        const R cz = d1[0] * d2[1] - d1[1] * d2[0];
        return std::sqrt((cx * cx + cy * cy) + cz * cz) / ldp3;
    }

    // copyParticleInfo tpp:997-1049 (positions = front after the step; curvature only when
    // Time > 0; expressions and volumes at every save)
    void save(Outputs<R> & out, bool time_positive)
    {
        for (size_t i = 0; i < n_; ++i)
            for (int d = 0; d < D_; ++d) out.positions.push_back(pos_[i * 3 + d]);
        out.cols++;
        if (s_.curvature == 3 && time_positive)
        {
            out.curvature.insert(out.curvature.end(), curv_.begin(), curv_.end());
            out.curv_cols++;
        }
        if (s_.nexpr > 0) out.exprs.insert(out.exprs.end(), expr_.begin(), expr_.end());
        if (s_.nvol > 0) out.volumes.insert(out.volumes.end(), vol_.begin(), vol_.end());
    }
};

} // namespace oracle
