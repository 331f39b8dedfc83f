// ORACLE — test infrastructure only (see granite_oracle.hpp). C entry points for ctypes.
//
// oracle_run() integrates a whole job on the CPU in FP32 (real64 = 0, the plain FP32 CPU path)
// or FP64 (real64 = 1, the FP64 oracle) and keeps the results in a handle; oracle_get() copies
// one output (converted to double) into caller memory; oracle_free() releases the handle.
#include <cstdio>
#include <cstring>
#include <memory>

#include "granite_oracle.hpp"

extern "C" {

struct OracleArgs
{
    int32_t real64;
    int32_t dim, integrator, border, curvature, abort_mode;
    int32_t uplift_mode, windfield_mode, constants_mode, volume_mode;
    int32_t reverse, weight_mode, time64, kahan;
    float sphere_radius, dt, dtd, max_length, max_sim_time;
    float grid_scale[3];
    float offset[3];
    uint64_t save_interval;
    int32_t min_alive, nx, ny, nz, nslices;
    int32_t nexpr, nconst, nvol, nthreads, allow_abs;
    const char * exprs[4];
    const char * const_names[4];
    const char * vol_names[8];
    const float * wind;
    const float * topo;
    const float * uplift;
    const float * constants;
    const float * volumes;
    const float * seeds;
    uint64_t n;
    const float * zlev; // stretched z axis (nz levels) or NULL
};

enum OracleOut
{
    O_POSITIONS = 0,
    O_CURVATURE = 1,
    O_EXPRS = 2,
    O_VOLUMES = 3,
    O_STATUS = 4,
    O_ABORT_STEP = 5,
    O_TOTAL_CURVATURE = 6,
    O_LENGTHS = 7
};

struct OracleResult
{
    std::vector<double> positions, curvature, exprs, volumes, total_curvature, lengths;
    std::vector<int32_t> status;
    std::vector<int64_t> abort_step;
    uint64_t steps{0}, cols{0}, curv_cols{0};
};

} // extern "C"

template <class R>
static void collect(const oracle::Outputs<R> & o, OracleResult & r)
{
    r.positions.assign(o.positions.begin(), o.positions.end());
    r.curvature.assign(o.curvature.begin(), o.curvature.end());
    r.exprs.assign(o.exprs.begin(), o.exprs.end());
    r.volumes.assign(o.volumes.begin(), o.volumes.end());
    r.total_curvature.assign(o.total_curvature.begin(), o.total_curvature.end());
    r.lengths.assign(o.lengths.begin(), o.lengths.end());
    r.status = o.status;
    r.abort_step = o.abort_step;
    r.steps = o.steps;
    r.cols = o.cols;
    r.curv_cols = o.curv_cols;
}

extern "C" {

void * oracle_run(const OracleArgs * a, char * err, int errlen)
{
    try
    {
        oracle::Settings s;
        s.dim = a->dim;
        s.integrator = a->integrator;
        s.border = a->border;
        s.curvature = a->curvature;
        s.abort_mode = a->abort_mode;
        s.uplift_mode = a->uplift_mode;
        s.windfield_mode = a->windfield_mode;
        s.constants_mode = a->constants_mode;
        s.volume_mode = a->volume_mode;
        s.reverse = a->reverse;
        s.weight_mode = a->weight_mode;
        s.time64 = a->time64;
        s.kahan = a->kahan;
        s.sphere_radius = a->sphere_radius;
        s.dt = a->dt;
        s.dtd = a->dtd;
        s.max_length = a->max_length;
        s.max_sim_time = a->max_sim_time;
        for (int d = 0; d < 3; ++d)
        {
            s.grid_scale[d] = a->grid_scale[d];
            s.offset[d] = a->offset[d];
        }
        s.save_interval = a->save_interval;
        s.min_alive = a->min_alive;
        s.nx = a->nx;
        s.ny = a->ny;
        s.nz = a->dim == 3 ? a->nz : 1;
        s.nslices = a->nslices;
        s.nconst = a->constants ? a->nconst : 0;
        s.nvol = a->volumes ? a->nvol : 0;
        s.have_topo = a->topo != nullptr;
        if (a->zlev && a->dim == 3) s.zlev.assign(a->zlev, a->zlev + s.nz);

        // identifier environment: volume keys -> F0.., constant keys -> C0.. (cpp:155-226)
        std::map<std::string, int> env;
        for (int k = 0; k < s.nvol; ++k) env[a->vol_names[k]] = oracle::N_F0 + k;
        for (int k = 0; k < s.nconst; ++k) env[a->const_names[k]] = oracle::N_C0 + k;
        s.nexpr = a->nexpr;
        for (int e = 0; e < a->nexpr; ++e)
        {
            std::string src(a->exprs[e]);
            oracle::Parser P(src, env, a->allow_abs != 0);
            oracle::Program prog = P.parse();
            // validation tpp:661-696
            if (s.dim == 2 && oracle::contains(prog, oracle::N_Z))
                throw std::runtime_error("Compute code contains undefinded token 'z'");
            if (s.dim == 2 && oracle::contains(prog, oracle::N_W))
                throw std::runtime_error("Compute code contains undefinded token 'w'");
            if (s.curvature != 3 && oracle::contains(prog, oracle::N_CV))
                throw std::runtime_error("Compute code contains undefinded token 'cv'");
            s.exprs.push_back(std::move(prog));
        }

        oracle::Inputs in;
        in.wind = a->wind;
        in.topo = a->topo;
        in.uplift = a->uplift;
        in.constants = a->constants;
        in.volumes = a->volumes;

        auto r = std::make_unique<OracleResult>();
        const int nthreads = a->nthreads > 0 ? a->nthreads : 1;
        if (a->real64)
        {
            oracle::Integrator<double> I(s, in, a->seeds, a->n);
            collect(I.run(nthreads), *r);
        }
        else
        {
            oracle::Integrator<float> I(s, in, a->seeds, a->n);
            collect(I.run(nthreads), *r);
        }
        return r.release();
    }
    catch (const std::exception & e)
    {
        if (err && errlen > 0) std::snprintf(err, size_t(errlen), "%s", e.what());
        return nullptr;
    }
}

void oracle_counts(const void * h, uint64_t * steps, uint64_t * cols, uint64_t * curv_cols)
{
    const OracleResult * r = static_cast<const OracleResult *>(h);
    *steps = r->steps;
    *cols = r->cols;
    *curv_cols = r->curv_cols;
}

// Copies output `which` into dst; returns the number of elements, or -1. With dst == NULL
// only the element count is returned.
int64_t oracle_get(const void * h, int32_t which, void * dst)
{
    const OracleResult * r = static_cast<const OracleResult *>(h);
    const std::vector<double> * v = nullptr;
    switch (which)
    {
        case O_POSITIONS: v = &r->positions; break;
        case O_CURVATURE: v = &r->curvature; break;
        case O_EXPRS: v = &r->exprs; break;
        case O_VOLUMES: v = &r->volumes; break;
        case O_TOTAL_CURVATURE: v = &r->total_curvature; break;
        case O_LENGTHS: v = &r->lengths; break;
        case O_STATUS:
            if (dst) std::memcpy(dst, r->status.data(), r->status.size() * sizeof(int32_t));
            return int64_t(r->status.size());
        case O_ABORT_STEP:
            if (dst) std::memcpy(dst, r->abort_step.data(), r->abort_step.size() * sizeof(int64_t));
            return int64_t(r->abort_step.size());
        default: return -1;
    }
    if (dst) std::memcpy(dst, v->data(), v->size() * sizeof(double));
    return int64_t(v->size());
}

void oracle_free(void * h) { delete static_cast<OracleResult *>(h); }

// Expression check: parses `src` with the given environment and evaluates it on `ref` (19
// slots) in float and double. Returns 0 on success, -1 on a parse error (message in err).
int oracle_eval_expr(const char * src, int32_t nnames, const char * const * names,
                     const int32_t * types, int32_t allow_abs, const double * ref, float * out32,
                     double * out64, char * err, int errlen)
{
    try
    {
        std::map<std::string, int> env;
        for (int i = 0; i < nnames; ++i) env[names[i]] = types[i];
        std::string s(src);
        oracle::Parser P(s, env, allow_abs != 0);
        oracle::Program p = P.parse();
        float rf[19];
        for (int i = 0; i < 19; ++i) rf[i] = static_cast<float>(ref[i]);
        *out32 = oracle::eval<float>(p, p.root, rf);
        *out64 = oracle::eval<double>(p, p.root, ref);
        return 0;
    }
    catch (const std::exception & e)
    {
        if (err && errlen > 0) std::snprintf(err, size_t(errlen), "%s", e.what());
        return -1;
    }
}

} // extern "C"
