// `pygranite` Python module: drop-in for the reference's pybind11 API
// (source/pygranite.cpp:3-152) -- enums, IntegratorSettings, DataLoader (Python-subclassable),
// TrajectorySet and TrajectoryIntegrator(settings, loader, set).compute() -- implemented over
// the C ABI of libgranite_hip.so (include/granite_hip.h). This file holds no device code: the
// host schedule (windows, FP32 time, saves, loader streaming) runs here, every step runs in
// the HIP library.
//
// Differences from the reference, all deliberate:
//  * steps run in multi-step launches per window (not one launch + device sync per step);
//  * the next slice is uploaded on a copy stream while the current window computes;
//  * the loader is kept alive by the integrator (the reference holds a bare reference);
//  * compute() may be called once; a second call raises (the reference dereferences null);
//  * volume saves keep every key (the reference's copy loop skips every second key,
//    TrajectoryIntegrator.tpp:1035-1044);
//  * identifiers are bound in the loader dict's order (the reference uses unordered_map
//    iteration order);
//  * extension settings (storage type, Kahan, FP64 time, local coordinates, texture-weight
//    emulation, abs() in expressions) default to the reference's behaviour.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "granite_hip.h"

namespace py = pybind11;

namespace granite
{

// ---- enums: source/granite/IntegratorSettings.hpp:16-76, AbortReason.hpp:8-15 ------------
enum class Space : uint32_t { Space3D = 0, Space2D = 1 };
enum class Integrator : uint32_t { ExplicitEuler = 0, ClassicRungeKutta = 1 };
enum class BorderMode : uint32_t { Block = 0, LoopX = 1, LoopY = 2, LoopXY = 3 };
enum class CurvatureMode : uint32_t
{
    Off = 0,
    TotalCurvature = 1,
    FastTotalCurvature = 2,
    IndividualAndTotalCurvature = 3
};
enum class AbortMode : uint32_t { Time = 0, Length = 1, FitLength = 2 };
enum class UpLiftMode : uint32_t { Off = 0, Constant = 1, Dynamic = 2 };
enum class WindfieldMode : uint32_t { Constant = 0, Dynamic = 1 };
enum class ConstantsMode : uint32_t { Off = 0, Constant = 1, Dynamic = 2 };
enum class AdditionalVolumeMode : uint32_t { Off = 0, Constant = 1, Dynamic = 2 };
enum class AbortReason : int32_t { Time = 0, Length = 1, Domain = 2, Topography = 3, Wind = 4 };
// extensions
enum class StorageType : uint32_t { Float32 = 0, Float16 = 1, BFloat16 = 2 };
enum class TextureWeights : uint32_t { Exact = 0, Quantized8 = 1, Quantized8Truncated = 2 };

using farray = py::array_t<float, py::array::c_style | py::array::forcecast>;
using map_t = std::map<std::string, farray>;

[[noreturn]] static void user_error(const std::string & msg)
{
    throw std::runtime_error("ERROR < USER INPUT | " + msg + " >");
}

static void check(int rc)
{
    if (rc != GR_OK) throw std::runtime_error(std::string("ERROR < RUNTIME | ") + gr_last_error() + " >");
}

// ---- IntegratorSettings: source/granite/IntegratorSettings.hpp:83-182 ---------------------
struct IntegratorSettings
{
    float SphereRadius{0.f};
    granite::Space Space{granite::Space::Space3D};
    granite::Integrator Integrator{granite::Integrator::ClassicRungeKutta};
    granite::WindfieldMode WindfieldMode{granite::WindfieldMode::Dynamic};
    granite::ConstantsMode ConstantsMode{granite::ConstantsMode::Off};
    granite::AdditionalVolumeMode AdditionalVolumeMode{granite::AdditionalVolumeMode::Off};
    granite::UpLiftMode UpLiftMode{granite::UpLiftMode::Off};
    float DeltaT{1};
    float DataTimeDistance{60};
    std::vector<float> GridScale{25, 25, 25};
    std::vector<float> Offset{0, 0, 0};
    size_t SaveInterval{1};
    bool Reverse{false};
    granite::BorderMode BorderMode{granite::BorderMode::Block};
    granite::CurvatureMode CurvatureMode{granite::CurvatureMode::Off};
    float MaximumLength{0.f};
    float MaximumSimulationTime{100.f};
    granite::AbortMode AbortMode{granite::AbortMode::Time};
    int32_t MinimumAliveParticles{0};
    py::array_t<float> Topography = py::array_t<float>(0);
    std::vector<std::string> AdditionalCompute;
    // extensions (defaults = reference behaviour)
    granite::StorageType WindfieldStorage{granite::StorageType::Float32};
    granite::TextureWeights TextureWeights{granite::TextureWeights::Exact};
    bool KahanSummation{false};
    bool TimeFP64{false};
    bool LocalCoordinates{false};
    bool AllowAbs{false};
    uint32_t BlockSize{0};
    bool SortParticles{true}; // Morton-reorder particle state at window edges (results unchanged)
    // and also every RegroupInterval steps inside a window while at least RegroupMinLive particles
    // are live (0: window edges only); launches of many live particles gather faster when a
    // wavefront's particles are neighbours, and they drift apart within a window. Applies to runs
    // that save no columns (SaveInterval 0): with saves, compute() is bound by the column reads
    // and the extra chunks only add work (ABC, 2^18 particles, every step saved: 0.41 s with
    // in-window regrouping, 0.29 s without)
    uint32_t RegroupInterval{64};
    uint64_t RegroupMinLive{65536};
    // stretched z axis (3D): height of every z node relative to Offset.Z, strictly increasing
    // from 0; empty = uniform spacing GridScale[2]. The z extent of the domain is the top level.
    std::vector<float> ZLevels;
};

// ---- DataLoader: source/granite/DataLoader.hpp:26-144 ------------------------------------
class DataLoader
{
public:
    DataLoader() = default;
    virtual ~DataLoader() = default;
    virtual bool step() = 0;
    virtual std::vector<farray> windfield() = 0;
    virtual farray uplift() { return farray(); }
    virtual map_t constants() { return map_t(); }
    virtual map_t additionalVolumes() { return map_t(); }
};

class PyDataLoader : public DataLoader
{
public:
    using DataLoader::DataLoader;
    bool step() override { PYBIND11_OVERRIDE_PURE(bool, DataLoader, step); }
    std::vector<farray> windfield() override
    {
        PYBIND11_OVERRIDE_PURE(std::vector<farray>, DataLoader, windfield);
    }
    farray uplift() override { PYBIND11_OVERRIDE(farray, DataLoader, uplift); }
    map_t constants() override { PYBIND11_OVERRIDE(map_t, DataLoader, constants); }
    map_t additionalVolumes() override { PYBIND11_OVERRIDE(map_t, DataLoader, additionalVolumes); }
};

// Saved columns are filled by the library's threaded copies straight from pinned memory: their
// storage is left uninitialised on allocation (default-initialised floats), so the first touch
// of the pages happens in those parallel copies rather than in a serial zero fill.
template <class T> struct NoInitAlloc : std::allocator<T>
{
    template <class U> struct rebind
    {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <class U> NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
    template <class U> void construct(U * p) noexcept { ::new (static_cast<void *>(p)) U; }
    template <class U, class... A> void construct(U * p, A &&... a)
    {
        ::new (static_cast<void *>(p)) U(std::forward<A>(a)...);
    }
};
using column_t = std::vector<float, NoInitAlloc<float>>;

// Runs f(begin, end) over [0, n) split across the host's threads (at most 16) when n is large.
template <class F> static void parallel_for(size_t n, size_t grain, F && f)
{
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>(std::min<size_t>(16, hw), (n + grain - 1) / std::max<size_t>(1, grain));
    if (nt <= 1)
    {
        f(size_t(0), n);
        return;
    }
    std::vector<std::thread> th;
    for (size_t t = 1; t < nt; ++t) th.emplace_back([&, t] { f(n * t / nt, n * (t + 1) / nt); });
    f(0, n / nt);
    for (auto & x : th) x.join();
}

// out[(i * S + s) * w + d] = cols[s][i * w + d]: columns of n particles (w floats each) to the
// particle-major layout of the exports, in tiles of particles x columns, on the host's threads.
static void transpose_columns(const std::vector<column_t> & cols, size_t n, size_t w, float * out)
{
    const size_t S = cols.size();
    parallel_for(n, 4096, [&](size_t b, size_t e) {
        for (size_t i0 = b; i0 < e; i0 += 64)
        {
            const size_t i1 = std::min(e, i0 + 64);
            for (size_t s0 = 0; s0 < S; s0 += 64)
            {
                const size_t s1 = std::min(S, s0 + 64);
                for (size_t s = s0; s < s1; ++s)
                {
                    const float * c = cols[s].data();
                    for (size_t i = i0; i < i1; ++i)
                        for (size_t d = 0; d < w; ++d) out[(i * S + s) * w + d] = c[i * w + d];
                }
            }
        }
    });
}

// ---- TrajectorySet: source/granite/TrajectorySet.{hpp,cpp} -------------------------------
class TrajectorySet
{
public:
    Space space{Space::Space2D};
    size_t num{0};
    std::vector<column_t> traj;   // per column N*D
    std::vector<column_t> curv;   // per column N
    std::vector<std::vector<column_t>> compute; // [program][column] N
    std::vector<std::string> vol_keys;
    std::vector<std::vector<column_t>> volumes; // [volume][column] N
    std::vector<float> total_curvature;
    std::vector<float> lengths;
    std::vector<int32_t> abort_reasons;
    std::vector<int32_t> abort_steps;
    size_t num_compute{0};
    double device_seconds{0.0}; // summed step-kernel time of the compute() that produced the set

    TrajectorySet() = default;
    TrajectorySet(Space s, size_t n) : space(s), num(n) {}

    explicit TrajectorySet(const farray & points)
    {
        if (points.ndim() != 2) user_error("Currently only 1 dimensional initial pointsets are supported.");
        num = size_t(points.shape(0));
        if (points.shape(1) == 2)
            space = Space::Space2D;
        else if (points.shape(1) == 3)
            space = Space::Space3D;
        else
            user_error("TrajectorySet does not support the given type of coordinates.");
        const float * p = points.data();
        traj.emplace_back(p, p + num * dim());
    }

    size_t dim() const { return space == Space::Space3D ? 3 : 2; }
    size_t lengthTrajectories() const { return traj.size(); }
    size_t numberTrajectories() const { return num; }

    py::list abortReasons() const
    {
        py::list l;
        for (int32_t r : abort_reasons) l.append(static_cast<AbortReason>(r));
        return l;
    }

    // (N, T, D): particle-major (TrajectorySet.cpp:143-160)
    py::array_t<float> trajectories() const
    {
        const size_t T = traj.size(), D = dim();
        py::array_t<float> r({py::ssize_t(num), py::ssize_t(T), py::ssize_t(D)});
        float * o = r.mutable_data();
        py::gil_scoped_release nogil;
        transpose_columns(traj, num, D, o);
        return r;
    }

    // (T*N, D), particle-major like trajectories() (TrajectorySet.cpp:131-141)
    py::array_t<float> cloud() const
    {
        py::array_t<float> t = trajectories();
        return t.reshape({py::ssize_t(traj.size() * num), py::ssize_t(dim())});
    }

    py::array_t<float> trajectory(size_t pid) const
    {
        if (pid >= num) throw py::index_error("particle index out of range");
        const size_t T = traj.size(), D = dim();
        py::array_t<float> r({py::ssize_t(T), py::ssize_t(D)});
        float * o = r.mutable_data();
        for (size_t t = 0; t < T; ++t)
            for (size_t d = 0; d < D; ++d) o[t * D + d] = traj[t][pid * D + d];
        return r;
    }

    static py::array_t<float> columns(const std::vector<column_t> & cols, size_t n)
    {
        py::array_t<float> r({py::ssize_t(n), py::ssize_t(cols.size())});
        float * o = r.mutable_data();
        py::gil_scoped_release nogil;
        transpose_columns(cols, n, 1, o);
        return r;
    }

    py::array_t<float> vec(const std::vector<float> & v) const
    {
        py::array_t<float> r(py::ssize_t(v.size()));
        if (!v.empty()) std::memcpy(r.mutable_data(), v.data(), v.size() * sizeof(float));
        return r;
    }

    py::array_t<float> totalCurvatures() const { return vec(total_curvature); }
    py::array_t<float> individualLengths() const { return vec(lengths); }
    py::array_t<float> curvatures() const { return columns(curv, num); }

    py::array_t<float> computeInfo(size_t id) const
    {
        if (id < num_compute && id < compute.size()) return columns(compute[id], num);
        user_error("TrajectorySet: Volume with given id does not exist!");
    }

    py::array_t<float> volumeInfo(const std::string & key) const
    {
        for (size_t k = 0; k < vol_keys.size(); ++k)
            if (vol_keys[k] == key) return columns(volumes[k], num);
        user_error("TrajectorySet: Volume with given key '" + key + "' does not exist!");
    }

    py::array_t<int32_t> abortSteps() const
    {
        py::array_t<int32_t> r(py::ssize_t(abort_steps.size()));
        if (!abort_steps.empty())
            std::memcpy(r.mutable_data(), abort_steps.data(), abort_steps.size() * sizeof(int32_t));
        return r;
    }
};

// ---- device context ------------------------------------------------------------------------
// Every TrajectoryIntegrator owns its own context (its own streams and scratch), so integrators
// used from several Python threads never share a context: the C ABI's one-submitting-thread rule
// holds per integrator, and their device work may overlap. The reference holds the GIL through
// compute() (source/pygranite.cpp:126-130); here the GIL is released around device work.
struct CtxHolder
{
    gr_ctx * c{nullptr};
    CtxHolder()
    {
        int n = 0;
        gr_device_count(&n);
        if (n == 0)
            throw std::runtime_error("ERROR < RUNTIME | pygranite_amd needs a HIP GPU (none visible) >");
        check(gr_ctx_create(0, &c));
        // one job per integrator: no later job would take a destroyed one's arrays
        check(gr_ctx_set_option(c, GR_CTX_POOL_BYTES, 0));
    }
    CtxHolder(const CtxHolder &) = delete;
    CtxHolder & operator=(const CtxHolder &) = delete;
    ~CtxHolder()
    {
        if (c) gr_ctx_destroy(c);
    }
};

struct FieldHolder
{
    gr_field * f{nullptr};
    FieldHolder() = default;
    FieldHolder(const FieldHolder &) = delete;
    FieldHolder & operator=(const FieldHolder &) = delete;
    ~FieldHolder()
    {
        if (f) gr_field_destroy(f);
    }
};

struct JobHolder
{
    gr_job * j{nullptr};
    ~JobHolder()
    {
        if (j) gr_job_destroy(j);
    }
};

// ---- TrajectoryIntegrator ------------------------------------------------------------------
class TrajectoryIntegrator
{
public:
    // TrajectoryIntegrator.hpp:138-147: initSet, initializeData, initializeTopography
    TrajectoryIntegrator(IntegratorSettings & settings, DataLoader & loader, TrajectorySet & set)
        : s_(settings), loader_(loader)
    {
        init_set(set);
        initialize_data();
        initialize_topography();
    }

    std::unique_ptr<TrajectorySet> compute();

private:
    IntegratorSettings s_;
    DataLoader & loader_;
    // declared before the fields so that it is destroyed after them (the context outlives its
    // fields and jobs)
    CtxHolder ctx_;
    std::unique_ptr<TrajectorySet> set_;
    std::vector<float> start_;
    uint32_t nx_{0}, ny_{0}, nz_{1};
    bool dims_loaded_{false};
    FieldHolder wind_, vols_, topo_;
    std::vector<std::string> vol_keys_, const_keys_;
    // per-particle data per slot (uplift N*D, constants N each), uploaded into the job
    std::vector<float> uplift_[3];
    std::vector<float> const_[3][GR_MAX_CONSTANTS];
    bool has_uplift_[3]{false, false, false};
    bool has_const_[3]{false, false, false};
    uint32_t back_{0}, front_{1}, cache_{2};
    bool computed_{false};

    bool dynamic_any() const
    {
        return s_.WindfieldMode == WindfieldMode::Dynamic || s_.UpLiftMode == UpLiftMode::Dynamic ||
               s_.AdditionalVolumeMode == AdditionalVolumeMode::Dynamic ||
               s_.ConstantsMode == ConstantsMode::Dynamic;
    }
    uint32_t D() const { return s_.Space == Space::Space3D ? 3 : 2; }

    // cpp:424-442
    void init_set(TrajectorySet & set)
    {
        set_ = std::make_unique<TrajectorySet>(set.space, set.num);
        if (set.traj.empty()) user_error("Provided TrajectorySet does not contain any data.");
        start_.assign(set.traj.back().begin(), set.traj.back().end());
        set_->traj.emplace_back(start_.begin(), start_.end());
        if (s_.Space != set.space)
            user_error("Dimensions of IntegratorSettings and provided TrajectorySet do not match.");
    }

    void check_grid(const farray & a)
    {
        const uint32_t d = D();
        if (a.ndim() != py::ssize_t(d))
            user_error(d == 3 ? "Invalid dataset was provided. Wrong amout of data (3D)."
                              : "Invalid dataset was provided. Wrong amout of data (2D).");
        uint32_t z = d == 3 ? uint32_t(a.shape(0)) : 1;
        uint32_t y = uint32_t(a.shape(d == 3 ? 1 : 0));
        uint32_t x = uint32_t(a.shape(d == 3 ? 2 : 1));
        if (!dims_loaded_)
        {
            nx_ = x;
            ny_ = y;
            nz_ = z;
            dims_loaded_ = true;
        }
        if (z != nz_) user_error("Invalid dataset was provided. Shape does not match in z dimension (3D).");
        if (y != ny_) user_error("Invalid dataset was provided. Shape does not match in y dimension.");
        if (x != nx_) user_error("Invalid dataset was provided. Shape does not match in x dimension.");
    }

    void load_wind(uint32_t slot)
    {
        std::vector<farray> w = loader_.windfield();
        const uint32_t d = D();
        if (w.size() < d) user_error("windfield() must return one array per velocity component.");
        for (uint32_t c = 0; c < d; ++c) check_grid(w[c]);
        if (!wind_.f)
        {
            gr_field_desc fd{};
            fd.nx = nx_;
            fd.ny = ny_;
            fd.nz = nz_;
            fd.ncomp = d;
            fd.dtype = uint32_t(s_.WindfieldStorage);
            fd.nslots = 3;
            fd.address = d == 3 ? 0 : 1; // 3D border, 2D clamp (cpp:503-506, 570-574)
            check(gr_field_create(ctx_.c, &fd, &wind_.f));
            if (d == 3 && !s_.ZLevels.empty())
            {
                if (s_.ZLevels.size() != nz_)
                    user_error("ZLevels needs one height per z node of the windfield.");
                check(gr_field_set_z_levels(wind_.f, s_.ZLevels.data(), uint32_t(s_.ZLevels.size())));
            }
        }
        const float * comps[3] = {w[0].data(), w[1].data(), d == 3 ? w[2].data() : nullptr};
        check(gr_field_upload(wind_.f, slot, comps, 1));
    }

    void load_volumes(uint32_t slot, bool initial)
    {
        map_t m = loader_.additionalVolumes();
        if (initial)
        {
            if (m.size() > GR_MAX_VOLUMES)
                py::print("WARNING < USER INPUT | You provided more additional volumes than supported. >");
            for (auto & kv : m)
            {
                if (vol_keys_.size() == GR_MAX_VOLUMES) break;
                vol_keys_.push_back(kv.first);
            }
            if (vol_keys_.empty()) return;
            gr_field_desc fd{};
            fd.nx = nx_;
            fd.ny = ny_;
            fd.nz = nz_;
            fd.ncomp = uint32_t(vol_keys_.size());
            fd.dtype = GR_F32;
            fd.nslots = 3;
            fd.address = D() == 3 ? 0 : 1;
            check(gr_field_create(ctx_.c, &fd, &vols_.f));
        }
        if (!vols_.f) return;
        std::vector<const float *> comps(vol_keys_.size(), nullptr);
        std::vector<farray> keep;
        for (size_t k = 0; k < vol_keys_.size(); ++k)
        {
            auto it = m.find(vol_keys_[k]);
            if (it == m.end()) user_error("additionalVolumes() is missing key '" + vol_keys_[k] + "'.");
            const farray & a = it->second;
            if (size_t(a.size()) != size_t(nx_) * ny_ * nz_)
                user_error("Additional volume '" + vol_keys_[k] + "' does not match the windfield grid.");
            keep.push_back(a);
            comps[k] = keep.back().data();
        }
        check(gr_field_upload(vols_.f, slot, comps.data(), 1));
    }

    void load_constants(uint32_t slot, bool initial)
    {
        map_t m = loader_.constants();
        if (m.size() > GR_MAX_CONSTANTS)
            py::print("WARNING < USER INPUT | You provided more constants than supported. >");
        if (initial)
            for (auto & kv : m)
            {
                if (const_keys_.size() == GR_MAX_CONSTANTS) break;
                const_keys_.push_back(kv.first);
            }
        for (size_t k = 0; k < const_keys_.size(); ++k)
        {
            auto it = m.find(const_keys_[k]);
            if (it == m.end()) user_error("constants() is missing key '" + const_keys_[k] + "'.");
            if (size_t(it->second.size()) < set_->num)
                user_error("Constant '" + const_keys_[k] + "' has fewer values than particles.");
            const float * p = it->second.data();
            const_[slot][k].assign(p, p + set_->num);
        }
        has_const_[slot] = true;
    }

    void load_uplift(uint32_t slot)
    {
        farray a = loader_.uplift();
        if (a.ndim() < 1 || size_t(a.shape(0)) != set_->num)
            user_error("Uplift dimension does not match number of particles.");
        if (size_t(a.size()) < set_->num * D()) user_error("Uplift must hold one vector per particle.");
        uplift_[slot].assign(a.data(), a.data() + set_->num * D());
        has_uplift_[slot] = true;
    }

    // cpp:146-259
    void initialize_data()
    {
        load_wind(0);
        if (s_.AdditionalVolumeMode != AdditionalVolumeMode::Off) load_volumes(0, true);
        if (s_.ConstantsMode != ConstantsMode::Off) load_constants(0, true);
        if (s_.UpLiftMode != UpLiftMode::Off) load_uplift(0);
        if (dynamic_any())
        {
            if (loader_.step())
                update_data(1);
            else
                user_error("Cannot increment loader.");
        }
    }

    // cpp:261-360
    void update_data(uint32_t slot)
    {
        if (s_.WindfieldMode == WindfieldMode::Dynamic) load_wind(slot);
        if (s_.AdditionalVolumeMode == AdditionalVolumeMode::Dynamic) load_volumes(slot, false);
        if (s_.ConstantsMode == ConstantsMode::Dynamic) load_constants(slot, false);
        if (s_.UpLiftMode == UpLiftMode::Dynamic) load_uplift(slot);
    }

    // cpp:635-663
    void initialize_topography()
    {
        if (s_.Topography.size() == 0) return;
        farray t = farray::ensure(s_.Topography);
        if (!t || t.ndim() != 2) user_error("Invalid topography was provided. Wrong dimension of data (2D).");
        if (uint32_t(t.shape(1)) != nx_ || uint32_t(t.shape(0)) != ny_)
            user_error("Topography dimensions does not match to provided data!");
        if (D() != 3) return; // 2D kernels never use topography (TrajectoryIntegratorMacros.hpp:71)
        gr_field_desc fd{};
        fd.nx = nx_;
        fd.ny = ny_;
        fd.nz = 1;
        fd.ncomp = 1;
        fd.dtype = GR_F32;
        fd.nslots = 1;
        fd.address = 1; // 2D texture (clamp)
        check(gr_field_create(ctx_.c, &fd, &topo_.f));
        const float * comps[1] = {t.data()};
        check(gr_field_upload(topo_.f, 0, comps, 0));
    }
};

std::unique_ptr<TrajectorySet> TrajectoryIntegrator::compute()
{
    if (computed_) throw std::runtime_error("ERROR < RUNTIME | compute() may only be called once per TrajectoryIntegrator >");
    computed_ = true;
    const uint32_t dim = D();
    const size_t n = set_->num;
    if (set_->traj.empty()) user_error("Provided TrajectorySet does not contain any data.");

    // kernel-variant availability of the reference (TrajectoryIntegrator.gen.py:52-53, 111-216)
    if (dim == 2 && (s_.CurvatureMode == CurvatureMode::TotalCurvature ||
                     s_.CurvatureMode == CurvatureMode::IndividualAndTotalCurvature))
        user_error("No kernel could be picked for your settings configuration.");
    if (s_.AdditionalCompute.empty() && s_.ConstantsMode != ConstantsMode::Off)
        user_error("No kernel could be picked for your settings configuration.");
    if (s_.AdditionalCompute.size() > GR_MAX_PROGRAMS)
        user_error("At most 4 AdditionalCompute programs are supported.");

    gr_job_desc d;
    std::memset(&d, 0, sizeof(d));
    d.dim = dim;
    d.integrator = uint32_t(s_.Integrator);
    d.border = dim == 3 ? 0 : uint32_t(s_.BorderMode); // 3D is always Block (Macros.hpp:32)
    d.curvature = uint32_t(s_.CurvatureMode);
    d.abort_mode = uint32_t(s_.AbortMode);
    d.uplift_mode = dim == 3 ? uint32_t(s_.UpLiftMode) : 0; // 2D: no uplift (Macros.hpp:67)
    d.windfield_mode = uint32_t(s_.WindfieldMode);
    d.constants_mode = uint32_t(s_.ConstantsMode);
    d.volume_mode = uint32_t(s_.AdditionalVolumeMode);
    d.reverse = s_.Reverse ? 1 : 0;
    d.weights = uint32_t(s_.TextureWeights);
    d.time64 = s_.TimeFP64 ? 1 : 0;
    d.kahan = s_.KahanSummation ? 1 : 0;
    d.local_coords = s_.LocalCoordinates ? 1 : 0;
    d.nvol = s_.AdditionalVolumeMode != AdditionalVolumeMode::Off ? uint32_t(vol_keys_.size()) : 0;
    d.nconst = s_.ConstantsMode != ConstantsMode::Off ? uint32_t(const_keys_.size()) : 0;
    d.dt = s_.DeltaT;
    d.dtd = s_.DataTimeDistance;
    d.sphere_radius = dim == 2 ? s_.SphereRadius : 0.f;
    d.max_length = s_.MaximumLength;

    // GridScale / Offset expansion (tpp:521-574)
    const auto & gs = s_.GridScale;
    const auto & of = s_.Offset;
    float g[3];
    if (gs.size() == 1)
        g[0] = g[1] = g[2] = gs[0];
    else if (gs.size() == 2)
    {
        g[0] = gs[0];
        g[1] = gs[1];
        g[2] = gs[0];
    }
    else if (gs.size() == 3)
    {
        g[0] = gs[0];
        g[1] = gs[1];
        g[2] = gs[2];
    }
    else
        user_error("Invalid GridScale dimension.");
    double o[3];
    if (of.size() == 1)
        o[0] = o[1] = o[2] = of[0];
    else if (gs.size() == 2 && of.size() >= 2)
    {
        o[0] = of[0];
        o[1] = of[1];
        o[2] = of[0];
    }
    else if (gs.size() == 3 && of.size() >= 3)
    {
        o[0] = of[0];
        o[1] = of[1];
        o[2] = of[2];
    }
    else
        user_error("Invalid Offset dimension.");
    for (int k = 0; k < 3; ++k)
    {
        d.grid_scale[k] = g[k];
        d.offset[k] = o[k];
    }
    d.domain[0] = float(nx_) * g[0];
    d.domain[1] = float(ny_) * g[1];
    d.domain[2] = float(nz_) * g[2];
    if (d.dim == 3 && !s_.ZLevels.empty()) d.domain[2] = s_.ZLevels.back();
    d.save_interval = s_.SaveInterval;

    // expressions (tpp:647-712)
    std::vector<const char *> cn, vn;
    for (auto & k : const_keys_) cn.push_back(k.c_str());
    for (auto & k : vol_keys_) vn.push_back(k.c_str());
    gr_expr_env env{};
    env.dim = dim;
    env.curvature = d.curvature;
    env.nconst = d.nconst;
    env.nvol = d.nvol;
    env.const_names = cn.empty() ? nullptr : cn.data();
    env.vol_names = vn.empty() ? nullptr : vn.data();
    env.flags = s_.AllowAbs ? GR_EXPR_ALLOW_ABS : 0;
    d.nprog = uint32_t(s_.AdditionalCompute.size());
    for (uint32_t e = 0; e < d.nprog; ++e)
        if (gr_expr_compile(s_.AdditionalCompute[e].c_str(), &env, &d.prog[e]) != GR_OK)
            user_error(gr_last_error());
    set_->num_compute = d.nprog;
    set_->compute.assign(d.nprog, {});
    set_->vol_keys = d.nvol ? vol_keys_ : std::vector<std::string>();
    set_->volumes.assign(d.nvol, {});

    // window plan: FP32 host time (tpp:778-995)
    const bool dyn = dynamic_any();
    const float dt = s_.DeltaT;
    const float max_steps_f = s_.MaximumSimulationTime / s_.DeltaT;
    const uint64_t SI = s_.SaveInterval;
    const size_t col_bytes = n * (dim + (d.curvature == 3 ? 1 : 0) + d.nprog + d.nvol) * sizeof(float);
    // device ring: enough for a window, bounded to ~1 GiB
    uint64_t per_window = SI ? uint64_t(std::ceil(double(s_.DataTimeDistance) / double(dt))) / SI + 2 : 1;
    if (!dyn && SI) per_window = uint64_t(std::max(0.f, max_steps_f)) / SI + 2;
    uint64_t cap = std::max<uint64_t>(1, std::min<uint64_t>(per_window, col_bytes ? (1ull << 30) / col_bytes : 1));
    cap = std::min<uint64_t>(cap, 4096);
    // the ring holds two chunks of saves: one being copied to the host while the next computes
    cap = std::max<uint64_t>(2, cap + (cap & 1));
    const uint64_t chunk_cols = cap / 2;

    JobHolder job;
    check(gr_job_create(ctx_.c, &d, start_.data(), n, uint32_t(cap), &job.j));
    if (s_.BlockSize) check(gr_job_set_launch(job.j, s_.BlockSize));
    for (uint32_t slot = 0; slot < 3; ++slot)
    {
        if (has_uplift_[slot] && d.uplift_mode) check(gr_job_set_particle_data(job.j, 0, slot, uplift_[slot].data()));
        if (has_const_[slot])
            for (uint32_t k = 0; k < d.nconst; ++k)
                check(gr_job_set_particle_data(job.j, 1 + k, slot, const_[slot][k].data()));
    }

    uint64_t counter = 0;
    float time32 = 0.f;
    double time64 = 0.0;
    float max_wt = s_.DataTimeDistance;
    double max_wt64 = s_.DataTimeDistance;
    uint32_t window = 0;
    uint64_t cols_read = 0;
    bool abort_due_alive = false;

    // Saved columns travel to the set asynchronously: the columns of a chunk are queued for
    // reading (device-to-host on the copy stream, straight into the set's column vectors) as soon
    // as the chunk is launched, and collected after the next chunk has been launched, so copies
    // overlap integration. Curvature has no column for the save of step 0 (Time == 0, tpp:1004).
    std::vector<uint64_t> pending;
    auto issue_reads = [&]() {
        uint64_t saved = 0;
        check(gr_job_saved_columns(job.j, &saved));
        const uint64_t cnt = saved - cols_read;
        if (!cnt) return;
        auto issue = [&](uint32_t kind, std::vector<column_t> & dst, uint64_t first, size_t width) {
            if (first >= saved) return;
            const uint64_t c = saved - first;
            std::vector<void *> cols(c);
            for (uint64_t k = 0; k < c; ++k)
            {
                dst.emplace_back(n * width);
                cols[k] = dst.back().data();
            }
            uint64_t t = 0;
            check(gr_job_read_saved_async(job.j, kind, first, c, cols.data(), &t));
            pending.push_back(t);
        };
        issue(GR_OUT_POSITIONS, set_->traj, cols_read, dim);
        if (d.curvature == GR_CURV_INDIVIDUAL_AND_TOTAL) issue(GR_OUT_CURVATURE, set_->curv, std::max<uint64_t>(cols_read, 1), 1);
        for (uint32_t e = 0; e < d.nprog; ++e) issue(GR_OUT_EXPR0 + e, set_->compute[e], cols_read, 1);
        for (uint32_t k = 0; k < d.nvol; ++k) issue(GR_OUT_VOLUME0 + k, set_->volumes[k], cols_read, 1);
        cols_read = saved;
    };
    auto collect = [&](size_t keep) {
        // waits for all but the `keep` most recent reads
        if (pending.size() <= keep) return;
        const size_t m = pending.size() - keep;
        {
            py::gil_scoped_release nogil;
            for (size_t i = 0; i < m; ++i) check(gr_job_read_wait(job.j, pending[i]));
        }
        pending.erase(pending.begin(), pending.begin() + long(m));
    };
    size_t reads_last = 0; // reads issued for the latest chunk
    auto read_columns = [&]() {
        const size_t before = pending.size();
        issue_reads();
        reads_last = pending.size() - before;
        collect(reads_last); // the previous chunk's columns are in the set now
    };

    auto alive_ok = [&]() {
        uint64_t a = 0;
        check(gr_job_count_alive(job.j, &a));
        const bool ok = int64_t(a) >= int64_t(s_.MinimumAliveParticles);
        if (!ok) abort_due_alive = true;
        return ok;
    };

    for (;;)
    {
        const float mwt = dyn ? max_wt : s_.MaximumSimulationTime;
        const double mwt64 = dyn ? max_wt64 : double(s_.MaximumSimulationTime);
        gr_window w;
        std::memset(&w, 0, sizeof(w));
        w.wind = wind_.f;
        w.wind_back = s_.WindfieldMode == WindfieldMode::Dynamic ? back_ : 0;
        w.wind_front = s_.WindfieldMode == WindfieldMode::Dynamic ? front_ : 0;
        w.volumes = d.nvol ? vols_.f : nullptr;
        w.vol_back = s_.AdditionalVolumeMode == AdditionalVolumeMode::Dynamic ? back_ : 0;
        w.vol_front = s_.AdditionalVolumeMode == AdditionalVolumeMode::Dynamic ? front_ : 0;
        w.topography = topo_.f;
        w.uplift_back = back_;
        w.uplift_front = front_;
        w.const_back = back_;
        w.const_front = front_;
        w.time_back = float(window) * s_.DataTimeDistance;
        w.time_front = float(window + 1) * s_.DataTimeDistance;
        w.time_back64 = double(window) * double(s_.DataTimeDistance);
        w.time_front64 = double(window + 1) * double(s_.DataTimeDistance);

        // particles moved during the last window: regroup them by cell (all saved columns
        // have been read at this point) while at least 4,096 are live -- fewer are stepped by a
        // handful of wavefronts whose time is their own step chain, not their gathers
        if (s_.SortParticles && n >= 4096)
        {
            uint64_t live = 0;
            check(gr_job_live_range(job.j, &live));
            if (live >= 4096) check(gr_job_reorder(job.j));
        }

        // run the window in chunks (each fits the column ring); with MinimumAliveParticles the
        // alive count gates every single step (tpp:973-976)
        bool window_done = false;
        while (!window_done)
        {
            uint32_t k = 0;
            float t = time32;
            double t64 = time64;
            uint64_t chunk_limit = s_.MinimumAliveParticles ? 1 : (SI ? chunk_cols * SI : (1u << 30));
            // chunks of RegroupInterval steps, each followed by a regrouping, while many
            // particles are live
            bool regroup = false;
            if (s_.SortParticles && s_.RegroupInterval && SI == 0 && n >= 4096 && !s_.MinimumAliveParticles)
            {
                uint64_t live = 0;
                check(gr_job_live_range(job.j, &live));
                if (live >= s_.RegroupMinLive)
                {
                    regroup = true;
                    chunk_limit = std::min<uint64_t>(chunk_limit, s_.RegroupInterval);
                }
            }
            while (uint64_t(k) < chunk_limit && float(counter + k) < max_steps_f)
            {
                if (s_.MinimumAliveParticles && !alive_ok()) break;
                k++;
                t += dt;
                t64 += double(dt);
                if (s_.TimeFP64 ? (t64 >= mwt64) : (t >= mwt))
                {
                    window_done = true;
                    break;
                }
                if (s_.MinimumAliveParticles) break;
            }
            if (k == 0)
            {
                window_done = true;
                break;
            }
            w.time_start = time32;
            w.time_start64 = time64;
            w.counter_start = counter;
            w.nsteps = k;
            check(gr_job_advance(job.j, &w));
            counter += k;
            time32 = t;
            time64 = t64;
            if (!window_done && !(float(counter) < max_steps_f)) window_done = true;
            // columns of earlier chunks are read while later chunks queue; the window's last
            // chunk is read after the loader has been stepped (overlapping the next upload)
            if (!window_done) read_columns();
            if (abort_due_alive) window_done = true;
            // every saved column has been read: regroup for the window's next chunk
            if (!window_done && regroup) check(gr_job_reorder(job.j));
        }

        bool stepped = false;
        if (dyn)
        {
            // loader.step() + upload into the cache slot overlap the queued window
            if (loader_.step())
            {
                update_data(cache_);
                stepped = true;
            }
        }
        read_columns();
        if (dyn)
        {
            if (!stepped) break;
            // upload per-particle data of the new cache slot into the job
            if (has_uplift_[cache_] && d.uplift_mode == 2) check(gr_job_set_particle_data(job.j, 0, cache_, uplift_[cache_].data()));
            if (d.constants_mode == 2 && has_const_[cache_])
                for (uint32_t k = 0; k < d.nconst; ++k)
                    check(gr_job_set_particle_data(job.j, 1 + k, cache_, const_[cache_][k].data()));
            if (abort_due_alive) break;
            max_wt += s_.DataTimeDistance;
            max_wt64 += double(s_.DataTimeDistance);
            const uint32_t tmp = back_;
            back_ = front_;
            front_ = cache_;
            cache_ = tmp;
            window++;
        }
        if (!(s_.WindfieldMode == WindfieldMode::Dynamic && float(counter) < max_steps_f)) break;
    }

    // final outputs (tpp:719-764, 914-915)
    collect(0);
    {
        py::gil_scoped_release nogil;
        check(gr_ctx_synchronize(ctx_.c));
        double ms = 0.0;
        check(gr_ctx_take_kernel_ms(ctx_.c, &ms, nullptr));
        set_->device_seconds = ms / 1e3;
    }
    if (SI == 0)
    {
        std::vector<float> pos(n * dim);
        check(gr_job_read(job.j, GR_OUT_POSITIONS, pos.data(), pos.size() * 4));
        set_->traj.emplace_back(pos.begin(), pos.end());
        if (d.curvature == GR_CURV_INDIVIDUAL_AND_TOTAL && time32 > 0.f)
        {
            std::vector<float> c(n);
            check(gr_job_read(job.j, GR_OUT_CURVATURE, c.data(), n * 4));
            set_->curv.emplace_back(c.begin(), c.end());
        }
        for (uint32_t e = 0; e < d.nprog; ++e)
        {
            std::vector<float> c(n);
            check(gr_job_read(job.j, GR_OUT_EXPR0 + e, c.data(), n * 4));
            set_->compute[e].emplace_back(c.begin(), c.end());
        }
        for (uint32_t k = 0; k < d.nvol; ++k)
        {
            std::vector<float> c(n);
            check(gr_job_read(job.j, GR_OUT_VOLUME0 + k, c.data(), n * 4));
            set_->volumes[k].emplace_back(c.begin(), c.end());
        }
    }
    set_->abort_reasons.resize(n);
    set_->abort_steps.resize(n);
    if (n)
    {
        check(gr_job_read(job.j, GR_OUT_STATUS, set_->abort_reasons.data(), n * 4));
        check(gr_job_read(job.j, GR_OUT_ABORT_STEP, set_->abort_steps.data(), n * 4));
    }
    if (d.abort_mode != GR_ABORT_TIME)
    {
        set_->lengths.resize(n);
        if (n) check(gr_job_read(job.j, GR_OUT_LENGTHS, set_->lengths.data(), n * 4));
    }
    if (d.curvature != GR_CURV_OFF)
    {
        set_->total_curvature.resize(n);
        if (n) check(gr_job_read(job.j, GR_OUT_TOTAL_CURVATURE, set_->total_curvature.data(), n * 4));
    }
    return std::move(set_);
}

} // namespace granite

PYBIND11_MODULE(pygranite, m)
{
    using namespace granite;
    m.doc() = "pygranite-compatible trajectory integration of massless particles in gridded, "
              "time-dependent 2D/3D wind fields on AMD Instinct MI355X (gfx950) GPUs.";

    py::enum_<Space>(m, "Space")
        .value("Space2D", Space::Space2D)
        .value("Space3D", Space::Space3D)
        .export_values();
    py::enum_<Integrator>(m, "Integrator")
        .value("ExplicitEuler", Integrator::ExplicitEuler)
        .value("ClassicRungeKutta", Integrator::ClassicRungeKutta)
        .export_values();
    py::enum_<AbortReason>(m, "AbortReason")
        .value("Time", AbortReason::Time)
        .value("Domain", AbortReason::Domain)
        .value("Length", AbortReason::Length)
        .value("Topography", AbortReason::Topography)
        .value("Wind", AbortReason::Wind)
        .export_values();
    py::enum_<BorderMode>(m, "BorderMode")
        .value("Block", BorderMode::Block)
        .value("LoopX", BorderMode::LoopX)
        .value("LoopY", BorderMode::LoopY)
        .value("LoopXY", BorderMode::LoopXY)
        .export_values();
    py::enum_<CurvatureMode>(m, "CurvatureMode")
        .value("Off", CurvatureMode::Off)
        .value("FastTotalCurvature", CurvatureMode::FastTotalCurvature)
        .value("IndividualAndTotalCurvature", CurvatureMode::IndividualAndTotalCurvature)
        .value("TotalCurvature", CurvatureMode::TotalCurvature)
        .export_values();
    py::enum_<AbortMode>(m, "AbortMode")
        .value("Time", AbortMode::Time)
        .value("Length", AbortMode::Length)
        .value("FitLength", AbortMode::FitLength)
        .export_values();
    py::enum_<UpLiftMode>(m, "UpLiftMode")
        .value("Off", UpLiftMode::Off)
        .value("Constant", UpLiftMode::Constant)
        .value("Dynamic", UpLiftMode::Dynamic)
        .export_values();
    py::enum_<WindfieldMode>(m, "WindfieldMode")
        .value("Constant", WindfieldMode::Constant)
        .value("Dynamic", WindfieldMode::Dynamic)
        .export_values();
    py::enum_<AdditionalVolumeMode>(m, "AdditionalVolumeMode")
        .value("Off", AdditionalVolumeMode::Off)
        .value("Constant", AdditionalVolumeMode::Constant)
        .value("Dynamic", AdditionalVolumeMode::Dynamic)
        .export_values();
    py::enum_<ConstantsMode>(m, "ConstantsMode")
        .value("Off", ConstantsMode::Off)
        .value("Constant", ConstantsMode::Constant)
        .value("Dynamic", ConstantsMode::Dynamic)
        .export_values();
    py::enum_<StorageType>(m, "StorageType")
        .value("Float32", StorageType::Float32)
        .value("Float16", StorageType::Float16)
        .value("BFloat16", StorageType::BFloat16);
    py::enum_<TextureWeights>(m, "TextureWeights")
        .value("Exact", TextureWeights::Exact)
        .value("Quantized8", TextureWeights::Quantized8)
        .value("Quantized8Truncated", TextureWeights::Quantized8Truncated);

    py::class_<IntegratorSettings>(m, "IntegratorSettings")
        .def(py::init<>())
        .def_readwrite("Space", &IntegratorSettings::Space)
        .def_readwrite("Integrator", &IntegratorSettings::Integrator)
        .def_readwrite("DeltaT", &IntegratorSettings::DeltaT)
        .def_readwrite("SaveInterval", &IntegratorSettings::SaveInterval)
        .def_readwrite("MinimumAliveParticles", &IntegratorSettings::MinimumAliveParticles)
        .def_readwrite("GridScale", &IntegratorSettings::GridScale)
        .def_readwrite("DataTimeDistance", &IntegratorSettings::DataTimeDistance)
        .def_readwrite("Topography", &IntegratorSettings::Topography)
        .def_readwrite("MaximumLength", &IntegratorSettings::MaximumLength)
        .def_readwrite("AdditionalCompute", &IntegratorSettings::AdditionalCompute)
        .def_readwrite("Offset", &IntegratorSettings::Offset)
        .def_readwrite("Reverse", &IntegratorSettings::Reverse)
        .def_readwrite("SphereRadius", &IntegratorSettings::SphereRadius)
        .def_readwrite("BorderMode", &IntegratorSettings::BorderMode)
        .def_readwrite("CurvatureMode", &IntegratorSettings::CurvatureMode)
        .def_readwrite("AbortMode", &IntegratorSettings::AbortMode)
        .def_readwrite("UpLiftMode", &IntegratorSettings::UpLiftMode)
        .def_readwrite("WindfieldMode", &IntegratorSettings::WindfieldMode)
        .def_readwrite("AdditionalVolumeMode", &IntegratorSettings::AdditionalVolumeMode)
        .def_readwrite("ConstantsMode", &IntegratorSettings::ConstantsMode)
        .def_readwrite("MaximumSimulationTime", &IntegratorSettings::MaximumSimulationTime)
        .def_readwrite("WindfieldStorage", &IntegratorSettings::WindfieldStorage)
        .def_readwrite("TextureWeights", &IntegratorSettings::TextureWeights)
        .def_readwrite("KahanSummation", &IntegratorSettings::KahanSummation)
        .def_readwrite("TimeFP64", &IntegratorSettings::TimeFP64)
        .def_readwrite("LocalCoordinates", &IntegratorSettings::LocalCoordinates)
        .def_readwrite("AllowAbs", &IntegratorSettings::AllowAbs)
        .def_readwrite("BlockSize", &IntegratorSettings::BlockSize)
        .def_readwrite("SortParticles", &IntegratorSettings::SortParticles)
        .def_readwrite("RegroupInterval", &IntegratorSettings::RegroupInterval)
        .def_readwrite("RegroupMinLive", &IntegratorSettings::RegroupMinLive)
        .def_readwrite("ZLevels", &IntegratorSettings::ZLevels);

    py::class_<TrajectoryIntegrator>(m, "TrajectoryIntegrator")
        .def(py::init<IntegratorSettings &, DataLoader &, TrajectorySet &>(), py::keep_alive<1, 3>())
        .def("compute", &TrajectoryIntegrator::compute);

    py::class_<DataLoader, PyDataLoader>(m, "DataLoader")
        .def(py::init<>())
        .def("step", &DataLoader::step)
        .def("windfield", &DataLoader::windfield)
        .def("uplift", &DataLoader::uplift)
        .def("constants", &DataLoader::constants)
        .def("additionalVolumes", &DataLoader::additionalVolumes);

    py::class_<TrajectorySet>(m, "TrajectorySet")
        .def(py::init<const farray &>())
        .def("trajectory", &TrajectorySet::trajectory)
        .def("trajectories", &TrajectorySet::trajectories)
        .def("cloud", &TrajectorySet::cloud)
        .def("totalCurvatures", &TrajectorySet::totalCurvatures)
        .def("curvatures", &TrajectorySet::curvatures)
        .def("individualLengths", &TrajectorySet::individualLengths)
        .def("volumeInfo", &TrajectorySet::volumeInfo)
        .def("computeInfo", &TrajectorySet::computeInfo)
        .def("lengthTrajectories", &TrajectorySet::lengthTrajectories)
        .def("abortReasons", &TrajectorySet::abortReasons)
        .def("numberTrajectories", &TrajectorySet::numberTrajectories)
        .def("abortSteps", &TrajectorySet::abortSteps)
        .def("deviceTime", [](const TrajectorySet & t) { return t.device_seconds; },
             "Summed device time in seconds of the step kernels of the compute() that produced "
             "this set (extension; 0 for a set built from an array).");

    m.attr("ABI_VERSION") = gr_abi_version();
}
