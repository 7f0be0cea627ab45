// batch.hip -- launch shapes, the autotuner and graph-captured batches of sets.
//
// Launch shape of a set: rows per workgroup of the stencil scatter and the order of the
// cut-element workgroups.  Neither changes a single value (each row and each cut element is
// computed by one thread / one wavefront in a fixed order), so the autotuner picks by time alone.
//
// A batch captures one assemble call of each of its sets (the same launch sequence as
// mgpu_set_assemble) into one HIP graph and replays it with one hipGraphLaunch.  The sets are
// spread round-robin over parallel branches forked from and joined into the capture stream, so
// that the small kernels of different sets (network, fit, cut elements) overlap on the device.
// Nothing in the captured sequence allocates or clears memory with hipMemset.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

#include "set_state.hpp"

using namespace mgpu;

namespace {

using TuneKey = std::tuple<int, int, int, int, int, int, int>;
std::mutex g_tune_mu;
std::map<TuneKey, std::pair<int, int>> g_tune_cache;

TuneKey tune_key(const mgpu_set* S) {
    const int n_cut = (int)S->h.cut_elems.size();
    const int bucket = n_cut == 0 ? 0 : 32 - __builtin_clz((unsigned)n_cut);
    return TuneKey{S->h.dim, S->h.nf, S->h.n_rows, bucket, S->precision, S->quadrature, S->h.generic ? 1 : 0};
}

// Generic sets have one launch shape per element order (no stencil); the order applies to the
// gather mode's element workgroups (scatter mode orders within each colour itself).
bool order_choice(const mgpu_set* S) {
    if (S->h.generic) return !S->h.scatter && !S->h.cut_elems.empty() && S->d_gen_order;
    return S->v.n_cut > 0;
}

void apply_tuning(mgpu_set* S, int variant, int order) {
    S->stencil_variant = variant;
    S->cut_order = order;
    S->v.cut_order = order ? S->d_cut_order : nullptr;
    S->g.order = order && S->h.generic && !S->h.scatter ? S->d_gen_order : nullptr;
}

}  // namespace

extern "C" mgpu_status mgpu_set_tuning_variants(const mgpu_set* set, int32_t* n_variants, int32_t* rows) {
    if (!set) return set_error(MGPU_ERR_INVALID_ARG, "null set");
    int n = 0;
    stencil_shape(set->h.nf, 0, &n);
    for (int v = 0; v < n; ++v) {
        int nv = 0;
        const int r = stencil_shape(set->h.nf, v, &nv);
        if (rows) rows[v] = r;
    }
    if (n_variants) *n_variants = n;
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_set_tuning(mgpu_set* set, int32_t stencil_variant, int32_t cut_order) {
    if (!set) return set_error(MGPU_ERR_INVALID_ARG, "null set");
    int nv = 0;
    stencil_shape(set->h.nf, 0, &nv);
    if (stencil_variant < -1 || stencil_variant >= nv) return set_error(MGPU_ERR_INVALID_ARG, "stencil variant out of range");
    if (cut_order != 0 && cut_order != 1) return set_error(MGPU_ERR_INVALID_ARG, "cut order must be 0 or 1");
    apply_tuning(set, stencil_variant, cut_order);
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_set_get_tuning(const mgpu_set* set, int32_t* stencil_variant, int32_t* cut_order) {
    if (!set) return set_error(MGPU_ERR_INVALID_ARG, "null set");
    if (stencil_variant) *stencil_variant = set->stencil_variant;
    if (cut_order) *cut_order = set->cut_order;
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_set_autotune(mgpu_set* set, const double* u, double* vals, double* rhs, int32_t iters,
                                         int32_t flags, double* ms) {
    MGPU_GUARD_BEGIN
    if (!set || iters < 1) return set_error(MGPU_ERR_INVALID_ARG, "bad argument");
    if (set->quadrature == MGPU_QUAD_LEARNED && !set->mlp) return set_error(MGPU_ERR_STATE, "no quadrature model");
    int nv = 0;
    stencil_shape(set->h.nf, 0, &nv);
    if (ms)
        for (int k = 0; k < 2 * nv; ++k) ms[k] = -1.0;
    if (set->h.n_rows == 0) return MGPU_OK;
    if (set->h.generic) nv = 1;   // shapes of a generic set: the element order only
    if (!u || !vals || !rhs) return set_error(MGPU_ERR_INVALID_ARG, "null buffer");
    const TuneKey key = tune_key(set);
    if (!(flags & 1)) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            apply_tuning(set, it->second.first, it->second.second);
            return MGPU_OK;
        }
    }
    hipEvent_t t0, t1;
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    double best = std::numeric_limits<double>::infinity();
    int best_v = 1, best_o = 0;
    mgpu_status rc = MGPU_OK;
    for (int v = 0; v < nv && rc == MGPU_OK; ++v)
        for (int o = 0; o < 2 && rc == MGPU_OK; ++o) {
            if (o == 1 && !order_choice(set)) continue;
            apply_tuning(set, v, o);
            if ((rc = launch_assemble(set, u, vals, rhs, nullptr)) != MGPU_OK) break;
            if (hipEventRecord(t0, set->stream) != hipSuccess) { rc = set_error(MGPU_ERR_HIP, "event record"); break; }
            for (int it = 0; it < iters && rc == MGPU_OK; ++it) rc = launch_assemble(set, u, vals, rhs, nullptr);
            if (rc != MGPU_OK) break;
            float t = 0.f;
            if (hipEventRecord(t1, set->stream) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
                hipEventElapsedTime(&t, t0, t1) != hipSuccess) {
                rc = set_error(MGPU_ERR_HIP, "event timing");
                break;
            }
            const double per = (double)t / iters;
            if (ms) ms[2 * v + o] = per;
            if (per < best) {
                best = per;
                best_v = v;
                best_o = o;
            }
        }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    apply_tuning(set, best_v, best_o);
    if (rc != MGPU_OK) return rc;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = {best_v, best_o};
    }
    return check_err_flag(set);
    MGPU_GUARD_END
}

// ------------------------------------------------------------------------------------------------
// batches
// ------------------------------------------------------------------------------------------------
struct mgpu_batch {
    std::vector<mgpu_set*> sets;
    std::vector<const double*> u;
    std::vector<double*> vals, rhs;
    hipStream_t stream = nullptr;          // replay stream (caller's)
    hipStream_t cap = nullptr;             // private capture stream
    std::vector<hipStream_t> branch;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    int n_branches = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void drop_graph(mgpu_batch* B) {
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    B->exec = nullptr;
    B->graph = nullptr;
}

mgpu_status capture(mgpu_batch* B, int nb) {
    drop_graph(B);
    while ((int)B->branch.size() < nb) {
        hipStream_t s;
        HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        B->branch.push_back(s);
        hipEvent_t e;
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        B->join.push_back(e);
    }
    HIP_TRY(hipStreamBeginCapture(B->cap, hipStreamCaptureModeThreadLocal));
    mgpu_status rc = MGPU_OK;
    hipError_t e = hipEventRecord(B->fork, B->cap);
    for (int k = 0; k < nb && e == hipSuccess; ++k) e = hipStreamWaitEvent(B->branch[k], B->fork, 0);
    for (size_t i = 0; i < B->sets.size() && e == hipSuccess && rc == MGPU_OK; ++i) {
        mgpu_set* S = B->sets[i];
        hipStream_t saved = S->stream;
        S->stream = B->branch[i % nb];
        rc = launch_assemble(S, B->u[i], B->vals[i], B->rhs[i], nullptr);
        S->stream = saved;
    }
    for (int k = 0; k < nb && e == hipSuccess; ++k) {
        e = hipEventRecord(B->join[k], B->branch[k]);
        if (e == hipSuccess) e = hipStreamWaitEvent(B->cap, B->join[k], 0);
    }
    hipGraph_t g = nullptr;
    const hipError_t e_end = hipStreamEndCapture(B->cap, &g);
    if (rc != MGPU_OK || e != hipSuccess || e_end != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (rc != MGPU_OK) return rc;
        return set_error(MGPU_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e != hipSuccess ? e : e_end));
    }
    B->graph = g;
    HIP_TRY(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
    B->n_branches = nb;
    return MGPU_OK;
}

mgpu_status time_replays(mgpu_batch* B, int iters, double* per) {
    hipEvent_t t0, t1;
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    hipError_t e = hipEventRecord(t0, B->stream);
    for (int it = 0; it < iters && e == hipSuccess; ++it) e = hipGraphLaunch(B->exec, B->stream);
    if (e == hipSuccess) e = hipEventRecord(t1, B->stream);
    if (e == hipSuccess) e = hipEventSynchronize(t1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (e != hipSuccess) return set_error(MGPU_ERR_HIP, std::string("batch timing: ") + hipGetErrorString(e));
    *per = (double)t / iters;
    return MGPU_OK;
}

}  // namespace

extern "C" mgpu_status mgpu_batch_destroy(mgpu_batch* B) {
    if (!B) return MGPU_OK;
    drop_graph(B);
    for (hipStream_t s : B->branch) (void)hipStreamDestroy(s);
    for (hipEvent_t e : B->join) (void)hipEventDestroy(e);
    if (B->fork) (void)hipEventDestroy(B->fork);
    if (B->cap) (void)hipStreamDestroy(B->cap);
    delete B;
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_batch_create(mgpu_set* const* sets, int32_t n_sets, const double* const* u,
                                         double* const* vals, double* const* rhs, int32_t n_branches,
                                         void* stream, mgpu_batch** out) {
    MGPU_GUARD_BEGIN
    if (!out || !sets || n_sets < 1 || !u || !vals || !rhs) return set_error(MGPU_ERR_INVALID_ARG, "bad argument");
    *out = nullptr;
    if (n_branches < 0 || n_branches > 64) return set_error(MGPU_ERR_INVALID_ARG, "n_branches must be 0..64");
    for (int i = 0; i < n_sets; ++i) {
        if (!sets[i]) return set_error(MGPU_ERR_INVALID_ARG, "null set in batch");
        if (sets[i]->h.n_rows > 0 && (!u[i] || !vals[i] || !rhs[i])) return set_error(MGPU_ERR_INVALID_ARG, "null buffer in batch");
        if (sets[i]->quadrature == MGPU_QUAD_LEARNED && !sets[i]->mlp) return set_error(MGPU_ERR_STATE, "no quadrature model");
        for (int j = 0; j < i; ++j)
            if (sets[j] == sets[i]) return set_error(MGPU_ERR_INVALID_ARG, "a set appears twice in the batch");
    }
    std::unique_ptr<mgpu_batch, mgpu_status (*)(mgpu_batch*)> B(new mgpu_batch, mgpu_batch_destroy);
    B->sets.assign(sets, sets + n_sets);
    B->u.assign(u, u + n_sets);
    B->vals.assign(vals, vals + n_sets);
    B->rhs.assign(rhs, rhs + n_sets);
    B->stream = static_cast<hipStream_t>(stream);
    HIP_TRY(hipStreamCreateWithFlags(&B->cap, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&B->fork, hipEventDisableTiming));
    // work of the sets' own streams that the graph must follow
    for (mgpu_set* S : B->sets) HIP_TRY(hipStreamSynchronize(S->stream));
    mgpu_status rc;
    if (n_branches > 0) {
        if ((rc = capture(B.get(), std::min(n_branches, n_sets))) != MGPU_OK) return rc;
    } else {
        // autotune the branch count on replays of the real graph
        double best = std::numeric_limits<double>::infinity();
        int best_nb = 1;
        for (int nb = 1; nb <= 8 && nb <= n_sets; nb *= 2) {
            if ((rc = capture(B.get(), nb)) != MGPU_OK) return rc;
            double per = 0.0;
            HIP_TRY(hipGraphLaunch(B->exec, B->stream));
            if ((rc = time_replays(B.get(), 3, &per)) != MGPU_OK) return rc;
            if (per < best) {
                best = per;
                best_nb = nb;
            }
        }
        if ((rc = capture(B.get(), best_nb)) != MGPU_OK) return rc;
    }
    *out = B.release();
    return MGPU_OK;
    MGPU_GUARD_END
}

extern "C" mgpu_status mgpu_batch_launch(mgpu_batch* B) {
    if (!B || !B->exec) return set_error(MGPU_ERR_INVALID_ARG, "null batch");
    HIP_TRY(hipGraphLaunch(B->exec, B->stream));
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_batch_synchronize(mgpu_batch* B) {
    if (!B) return set_error(MGPU_ERR_INVALID_ARG, "null batch");
    HIP_TRY(hipStreamSynchronize(B->stream));
    for (mgpu_set* S : B->sets) {
        mgpu_status rc = check_err_flag(S);
        if (rc != MGPU_OK) return rc;
    }
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_batch_time(mgpu_batch* B, int32_t iters, double* per_launch_ms, int32_t* n_branches) {
    MGPU_GUARD_BEGIN
    if (!B || !B->exec || iters < 1) return set_error(MGPU_ERR_INVALID_ARG, "bad argument");
    double per = 0.0;
    mgpu_status rc = time_replays(B, iters, &per);
    if (rc != MGPU_OK) return rc;
    if (per_launch_ms) *per_launch_ms = per;
    if (n_branches) *n_branches = B->n_branches;
    return mgpu_batch_synchronize(B);
    MGPU_GUARD_END
}
