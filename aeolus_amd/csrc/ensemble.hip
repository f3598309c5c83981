// Uneven-L ensembles: several member batches (one Gauss model per L class)
// advanced together.  Two RK4 steps of every group are captured once into a
// hipGraph (one branch per group, forked and joined with events so the
// classes run concurrently) and the graph is replayed: the host issues one
// launch per two steps for the whole ensemble.
#include <cmath>
#include <new>
#include <vector>

#include "plan.hpp"

struct aeo_ensemble_s {
    std::vector<aeo_model_s*> models;
    std::vector<hipStream_t> branch;   // one capture branch per group
    hipStream_t origin = nullptr;      // capture origin stream
    std::vector<hipEvent_t> ev;        // fork (0) and joins (1..ngroups)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double dt = NAN;
    // (mb, nslices, plan table generation) of every model at capture time
    std::vector<long long> cfg;
};

namespace aeo {

// One RK4 step of a Gauss model: internal state cur -> out (cur != out).
// The stage-0 synthesis of cur must already be in the model's E/O workspace;
// with syn_last the final stage leaves the synthesis of `out` there.
static int group_step(aeo_model_s* m, const double2* cur, double2* out, double dt, bool syn_last,
                      hipStream_t s) {
    const size_t st = (size_t)(m->L + 1) * (m->L + 1) * m->nstate * m->nmem;
    double2* ystage[2] = {m->d_ystage, m->d_ystage + st};
    for (int k = 0; k < 4; ++k) {
        const double2* ys = (k == 0) ? cur : ystage[(k - 1) % 2];
        double2* ynext = (k == 3) ? out : ystage[k % 2];
        const bool syn = (k < 3) || syn_last;
        int rc = gauss_model_stage(m, ys, cur, nullptr, nullptr, k, dt, ynext, syn, s);
        if (rc) return rc;
    }
    return AEO_OK;
}

// What a captured graph depends on beyond dt: the tiles (fused or GEMM form)
// and precision of every model and the generation of its plan's split tables
// (rebuilding them moves them, and the graph holds their addresses).
static std::vector<long long> graph_key(const aeo_ensemble_s* e) {
    std::vector<long long> k;
    for (const aeo_model_s* m : e->models) {
        k.push_back(m->mb);
        k.push_back(m->gemm_cg);
        k.push_back(m->gemm_me);
        k.push_back(m->gemm_ntw);
        k.push_back(m->nslices);
        k.push_back(m->plan->tables_gen);
    }
    return k;
}

static void release_graph(aeo_ensemble_s* e) {
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
}

static void free_ensemble(aeo_ensemble_s* e) {
    if (!e) return;
    release_graph(e);
    for (hipStream_t s : e->branch) (void)hipStreamDestroy(s);
    if (e->origin) (void)hipStreamDestroy(e->origin);
    for (hipEvent_t v : e->ev) (void)hipEventDestroy(v);
    delete e;
}

static int capture(aeo_ensemble_s* e, double dt) {
    release_graph(e);
    const int ng = (int)e->models.size();
    AEO_CHECK_HIP(hipStreamBeginCapture(e->origin, hipStreamCaptureModeThreadLocal));
    int rc = AEO_OK;
    hipError_t he = hipEventRecord(e->ev[0], e->origin);
    for (int g = 0; g < ng && rc == AEO_OK && he == hipSuccess; ++g) {
        aeo_model_s* m = e->models[g];
        const size_t st = (size_t)(m->L + 1) * (m->L + 1) * m->nstate * m->nmem;
        he = hipStreamWaitEvent(e->branch[g], e->ev[0], 0);
        if (he != hipSuccess) break;
        rc = group_step(m, m->d_ybuf, m->d_ybuf + st, dt, true, e->branch[g]);
        if (rc == AEO_OK) rc = group_step(m, m->d_ybuf + st, m->d_ybuf, dt, true, e->branch[g]);
        if (rc == AEO_OK) he = hipEventRecord(e->ev[g + 1], e->branch[g]);
        if (rc == AEO_OK && he == hipSuccess) he = hipStreamWaitEvent(e->origin, e->ev[g + 1], 0);
    }
    hipGraph_t graph = nullptr;
    hipError_t he_end = hipStreamEndCapture(e->origin, &graph);
    if (rc != AEO_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (he != hipSuccess || he_end != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        set_error(std::string("aeo_ensemble: graph capture failed: ") +
                  hipGetErrorString(he != hipSuccess ? he : he_end));
        return AEO_ERR_HIP;
    }
    e->graph = graph;
    AEO_CHECK_HIP(hipGraphInstantiate(&e->exec, graph, nullptr, nullptr, 0));
    e->dt = dt;
    e->cfg = graph_key(e);
    return AEO_OK;
}

}  // namespace aeo

using namespace aeo;

extern "C" {

int aeo_ensemble_create(int ngroups, const aeo_model_t* models, aeo_ensemble_t* out) {
    AEO_REQUIRE(out != nullptr, AEO_ERR_INVALID, "aeo_ensemble_create: out is NULL");
    *out = nullptr;
    AEO_REQUIRE(ngroups >= 1 && models != nullptr, AEO_ERR_INVALID,
                "aeo_ensemble_create: need ngroups >= 1 models");
    for (int g = 0; g < ngroups; ++g) {
        AEO_REQUIRE(models[g] != nullptr, AEO_ERR_INVALID, "aeo_ensemble_create: NULL model");
        AEO_REQUIRE(models[g]->plan->kind == kGauss, AEO_ERR_UNSUPPORTED,
                    "aeo_ensemble_create: ensembles need Gauss-plan models");
        for (int h = 0; h < g; ++h)
            AEO_REQUIRE(models[h] != models[g], AEO_ERR_INVALID,
                        "aeo_ensemble_create: a model appears twice");
    }
    aeo_ensemble_s* e = new (std::nothrow) aeo_ensemble_s();
    AEO_REQUIRE(e != nullptr, AEO_ERR_HIP, "aeo_ensemble_create: out of host memory");
    e->models.assign(models, models + ngroups);
    hipError_t he = hipStreamCreateWithFlags(&e->origin, hipStreamNonBlocking);
    for (int g = 0; g < ngroups && he == hipSuccess; ++g) {
        hipStream_t s = nullptr;
        he = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (he == hipSuccess) e->branch.push_back(s);
    }
    for (int i = 0; i <= ngroups && he == hipSuccess; ++i) {
        hipEvent_t v = nullptr;
        he = hipEventCreateWithFlags(&v, hipEventDisableTiming);
        if (he == hipSuccess) e->ev.push_back(v);
    }
    if (he != hipSuccess) {
        free_ensemble(e);
        set_error(std::string("aeo_ensemble_create: ") + hipGetErrorString(he));
        return AEO_ERR_HIP;
    }
    *out = e;
    return AEO_OK;
}

int aeo_ensemble_destroy(aeo_ensemble_t ens) {
    free_ensemble(ens);
    return AEO_OK;
}

int aeo_ensemble_rk4(aeo_ensemble_t ens, const double* const* y, double dt, int nsteps,
                     double* const* y_out, int use_graph, void* stream) {
    AEO_REQUIRE(ens && y && y_out, AEO_ERR_INVALID, "aeo_ensemble_rk4: NULL argument");
    AEO_REQUIRE(nsteps >= 0 && std::isfinite(dt), AEO_ERR_INVALID,
                "aeo_ensemble_rk4: need nsteps >= 0 and a finite dt");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int ng = (int)ens->models.size();
    for (int g = 0; g < ng; ++g)
        AEO_REQUIRE(y[g] && y_out[g], AEO_ERR_INVALID, "aeo_ensemble_rk4: NULL state pointer");
    // states in: boundary layout -> internal ybuf[0]; stage-0 synthesis
    for (int g = 0; g < ng; ++g) {
        aeo_model_s* m = ens->models[g];
        AEO_REQUIRE(m->prof.cls < 0, AEO_ERR_STATE,
                    "aeo_ensemble_rk4: disable aeo_model_profile on ensemble members");
        int rc = launch_transpose(reinterpret_cast<const double2*>(y[g]), m->d_ybuf, m->L + 1,
                                  m->nstate * m->nmem, s);
        if (!rc) rc = gauss_model_prime(m, m->d_ybuf, s);
        if (rc) return rc;
    }
    const int pairs = nsteps / 2;
    if (pairs > 0) {
        if (use_graph) {
            if (!ens->exec || ens->dt != dt || graph_key(ens) != ens->cfg) {
                // capture only records the launches; they execute in order on `s`
                int rc = capture(ens, dt);
                if (rc) return rc;
            }
            for (int i = 0; i < pairs; ++i) AEO_CHECK_HIP(hipGraphLaunch(ens->exec, s));
        } else {
            for (int i = 0; i < pairs; ++i)
                for (int g = 0; g < ng; ++g) {
                    aeo_model_s* m = ens->models[g];
                    const size_t st = (size_t)(m->L + 1) * (m->L + 1) * m->nstate * m->nmem;
                    int rc = group_step(m, m->d_ybuf, m->d_ybuf + st, dt, true, s);
                    if (!rc) rc = group_step(m, m->d_ybuf + st, m->d_ybuf, dt, true, s);
                    if (rc) return rc;
                }
        }
    }
    for (int g = 0; g < ng; ++g) {
        aeo_model_s* m = ens->models[g];
        const size_t st = (size_t)(m->L + 1) * (m->L + 1) * m->nstate * m->nmem;
        const double2* fin = m->d_ybuf;
        if (nsteps % 2) {
            int rc = group_step(m, m->d_ybuf, m->d_ybuf + st, dt, false, s);
            if (rc) return rc;
            fin = m->d_ybuf + st;
        }
        int rc = launch_transpose(fin, reinterpret_cast<double2*>(y_out[g]), m->L + 1,
                                  m->nstate * m->nmem, s);
        if (rc) return rc;
    }
    return AEO_OK;
}

}  // extern "C"
