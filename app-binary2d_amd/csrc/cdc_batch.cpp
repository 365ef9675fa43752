// cdc_batch.cpp -- one RK step of many independent runs per call (the
// ensemble path; no counterpart in the reference, whose scheme::advance
// steps one run, src/scheme.rs:323-349).
//
// Every member is an ordinary cdc_solver: the batch reads and writes the
// members' own device buffers, plans each member's stages with the same host
// code as cdc_solver_advance(member, 1), and launches each RK stage once for
// all members that share a cooling and viscosity law (the stage kernels are
// compiled for the law; finalize_kernel_batch once for all members), with
// the per-run arithmetic of the single-run launches: a member therefore ends
// bitwise where advancing it alone would have left it.  The whole step -- argument
// upload, flag reset, the stages, the per-member max signal speed and error
// flag read-back -- is captured once in a HIP graph and replayed while the
// set of active members stays the same.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/cdc_solver.h"
#include "cdc_host.h"
#include "cdc_internal.h"

using cdc::host::set_msg;
using cdc::host::StepPlan;

struct cdc_ensemble {
    std::vector<cdc_solver *> members;
    int nv = 0, rk = 0;
    cdc::TileShape ts{};
    hipStream_t stream = nullptr;
    bool use_graph = true;

    // per-step sizes of the captured sequence
    int cap_runs = 0;           // runs the buffers below hold
    cdc::StageArgs *h_args = nullptr;   // pinned, [stage][run]
    cdc::StageArgs *d_args = nullptr;
    int32_t *d_tile_begin = nullptr;   // per law group, each from 0: [nrun + groups]
    int32_t *d_block_begin = nullptr;  // [nrun + 1]
    double *d_amax = nullptr;
    int32_t *d_flags = nullptr;
    double *h_amax = nullptr;   // pinned
    int32_t *h_flags = nullptr; // pinned

    // host copy of each member's max signal speed, valid while the member's
    // committed state is the one this batch produced (state version)
    std::vector<double> amax;
    std::vector<uint64_t> amax_of;

    // the active runs of a step in launch order (member indexes, grouped by
    // law) and the law groups [r0, r1) of that order
    struct Group {
        int r0, r1, cool, visc, ntiles;
    };
    std::vector<Group> groups;

    // graph of the last active set, and the laws it was captured with
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<int> graph_runs, graph_laws;

    // statistics
    int64_t steps = 0;
    double cells = 0.0;
};

namespace {

void free_buffers(cdc_ensemble *e) {
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
    e->graph_runs.clear();
    e->graph_laws.clear();
    if (e->h_args) (void)hipHostFree(e->h_args);
    if (e->h_amax) (void)hipHostFree(e->h_amax);
    if (e->h_flags) (void)hipHostFree(e->h_flags);
    if (e->d_args) (void)hipFree(e->d_args);
    if (e->d_tile_begin) (void)hipFree(e->d_tile_begin);
    if (e->d_block_begin) (void)hipFree(e->d_block_begin);
    if (e->d_amax) (void)hipFree(e->d_amax);
    if (e->d_flags) (void)hipFree(e->d_flags);
    e->h_args = nullptr;
    e->h_amax = nullptr;
    e->h_flags = nullptr;
    e->d_args = nullptr;
    e->d_tile_begin = e->d_block_begin = nullptr;
    e->d_amax = nullptr;
    e->d_flags = nullptr;
    e->cap_runs = 0;
}

int alloc_buffers(cdc_ensemble *e, int n) {
    const size_t na = (size_t)e->rk * n;
    if (hipHostMalloc((void **)&e->h_args, na * sizeof(cdc::StageArgs), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&e->h_amax, n * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&e->h_flags, n * sizeof(int32_t), hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void **)&e->d_args, na * sizeof(cdc::StageArgs)) != hipSuccess ||
        hipMalloc((void **)&e->d_tile_begin, 2 * (size_t)n * sizeof(int32_t)) != hipSuccess ||
        hipMalloc((void **)&e->d_block_begin, (n + 1) * sizeof(int32_t)) != hipSuccess ||
        hipMalloc((void **)&e->d_amax, n * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&e->d_flags, n * sizeof(int32_t)) != hipSuccess) {
        (void)hipGetLastError();
        free_buffers(e);
        return set_msg(CDC_ERR_OUT_OF_MEMORY, "ensemble buffers");
    }
    e->cap_runs = n;
    return CDC_OK;
}

// The device work of one step for `runs` (indexes into members), in order.
int enqueue_step(cdc_ensemble *e, int nrun, int total_blocks) {
    const size_t na = (size_t)e->rk * nrun;
    HIP_TRY(hipMemcpyAsync(e->d_args, e->h_args, na * sizeof(cdc::StageArgs), hipMemcpyHostToDevice, e->stream));
    HIP_TRY(cdc::launch_batch_reset(e->d_args, nrun, e->stream));
    for (int k = 0; k < e->rk; ++k) {
        const cdc::StageArgs *a = e->d_args + (size_t)k * nrun;
        for (size_t g = 0; g < e->groups.size(); ++g) {
            const cdc_ensemble::Group &G = e->groups[g];
            HIP_TRY(cdc::launch_stage_batch(e->nv, e->ts, a + G.r0, e->d_tile_begin + G.r0 + g, G.r1 - G.r0,
                                            G.ntiles, nullptr, nullptr, 0, G.cool, G.visc, k > 0 ? 1 : 0, nullptr,
                                            nullptr, e->stream));
        }
        HIP_TRY(cdc::launch_finalize_batch(e->nv, a, e->d_block_begin, nrun, total_blocks, e->stream));
    }
    HIP_TRY(cdc::launch_batch_tail(e->nv, e->d_args + (size_t)(e->rk - 1) * nrun, nrun, e->d_amax, e->d_flags,
                                   e->stream));
    HIP_TRY(hipMemcpyAsync(e->h_amax, e->d_amax, nrun * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipMemcpyAsync(e->h_flags, e->d_flags, nrun * sizeof(int32_t), hipMemcpyDeviceToHost, e->stream));
    return CDC_OK;
}

}  // namespace

extern "C" {

int32_t cdc_ensemble_create(cdc_solver_t *const *members, int32_t n, cdc_ensemble_t **out) {
    if (!out) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    if (!members || n < 1) return set_msg(CDC_ERR_INVALID_ARGUMENT, "an ensemble needs at least one member");
    const cdc_solver *m0 = members[0];
    if (!m0) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null member");
    int bs_gcd = 0;
    for (int i = 0; i < n; ++i) {
        const cdc_solver *m = members[i];
        if (!m) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null member");
        if (m->nv != m0->nv) return set_msg(CDC_ERR_INVALID_ARGUMENT, "members must share one hydro kind");
        if (m->phys.rk_order != m0->phys.rk_order)
            return set_msg(CDC_ERR_INVALID_ARGUMENT, "members must share one rk_order");
        if (m->device != m0->device) return set_msg(CDC_ERR_INVALID_ARGUMENT, "members must share one device");
        for (int j = 0; j < i; ++j)
            if (members[j] == m) return set_msg(CDC_ERR_INVALID_ARGUMENT, "a member appears twice");
        int a = bs_gcd, b = m->bs;
        while (b) { int t = a % b; a = b; b = t; }
        bs_gcd = a;
    }
    // one tile shape for every member: the shape the members already share
    // when it divides every block size, else the first supported shape that
    // does (block sizes are multiples of 8)
    cdc::TileShape ts = m0->ts;
    bool shared = bs_gcd % ts.tx == 0 && bs_gcd % ts.ty == 0;
    for (int i = 1; i < n && shared; ++i)
        shared = members[i]->ts.tx == ts.tx && members[i]->ts.ty == ts.ty && members[i]->ts.nt == ts.nt;
    if (!shared) ts = cdc::tile_shape_for(bs_gcd, m0->nv);
    for (int i = 0; i < n; ++i) {
        const int rc = cdc_solver_set_kernel(members[i], ts.tx, ts.ty, ts.nt);
        if (rc) return rc;
    }
    cdc_ensemble *e = new cdc_ensemble();
    e->members.assign(members, members + n);
    e->nv = m0->nv;
    e->rk = m0->phys.rk_order;
    e->ts = ts;
    e->amax.assign(n, 0.0);
    e->amax_of.assign(n, UINT64_MAX);
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return set_msg(CDC_ERR_HIP, "hipStreamCreate failed");
    }
    const int rc = alloc_buffers(e, n);
    if (rc) {
        (void)hipStreamDestroy(e->stream);
        delete e;
        return rc;
    }
    *out = e;
    return CDC_OK;
}

void cdc_ensemble_destroy(cdc_ensemble_t *e) {
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    free_buffers(e);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

int32_t cdc_ensemble_set_graph(cdc_ensemble_t *e, int32_t enable) {
    if (!e) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null ensemble");
    e->use_graph = enable != 0;
    return CDC_OK;
}

int32_t cdc_ensemble_get_tile(cdc_ensemble_t *e, int32_t *tx, int32_t *ty) {
    if (!e) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null ensemble");
    if (tx) *tx = e->ts.tx;
    if (ty) *ty = e->ts.ty;
    return CDC_OK;
}

int32_t cdc_ensemble_get_stats(cdc_ensemble_t *e, int64_t *member_steps, double *zone_updates) {
    if (!e) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null ensemble");
    if (member_steps) *member_steps = e->steps;
    if (zone_updates) *zone_updates = e->cells;
    return CDC_OK;
}

int32_t cdc_ensemble_advance(cdc_ensemble_t *e, const int32_t *active, double *dt_out, int32_t *codes_out,
                             cdc_error_t *errs_out) {
    if (!e) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null ensemble");
    const int n = (int)e->members.size();
    if (errs_out) memset(errs_out, 0, (size_t)n * sizeof(cdc_error_t));
    if (codes_out)
        for (int i = 0; i < n; ++i) codes_out[i] = CDC_OK;

    // active members in launch order: grouped by their cooling and viscosity
    // laws (their physics may change between steps), member order within a
    // group; each group's stages launch the kernels a member's own advance
    // launches
    std::vector<int> runs, laws;
    for (int i = 0; i < n; ++i)
        if (!active || active[i]) runs.push_back(i);
    if (runs.empty()) return CDC_OK;
    const int nrun = (int)runs.size();
    auto law = [&](int i) {
        const cdc_solver *m = e->members[(size_t)i];
        const int cool = m->nv == 3 ? cdc::kCoolAny
                         : m->hydro.cooling_model == CDC_COOLING_BETA ? cdc::kCoolBetaExp
                                                                      : cdc::kCoolT4;
        return cool * 4 + (m->phys.viscosity_model == CDC_VISC_ALPHA ? cdc::kViscAlpha : cdc::kViscConst);
    };
    std::stable_sort(runs.begin(), runs.end(), [&](int a, int b) { return law(a) < law(b); });
    for (int r = 0; r < nrun; ++r) laws.push_back(law(runs[(size_t)r]));

    // their dt and stage plans
    std::vector<StepPlan> plans((size_t)nrun);
    std::vector<int32_t> tile_begin, block_begin((size_t)nrun + 1, 0);
    e->groups.clear();
    for (int r = 0; r < nrun; ++r) {
        cdc_solver *m = e->members[(size_t)runs[r]];
        if (m->ts.tx != e->ts.tx || m->ts.ty != e->ts.ty || m->ts.nt != e->ts.nt)
            return set_msg(CDC_ERR_INVALID_ARGUMENT, "a member's tile shape was changed after joining the ensemble");
        if (m->phys.rk_order != e->rk)
            return set_msg(CDC_ERR_INVALID_ARGUMENT, "a member's rk_order was changed after joining the ensemble");
        double amax = 0.0;
        if (e->amax_of[(size_t)runs[r]] == m->version && m->amax_valid) {
            amax = e->amax[(size_t)runs[r]];
        } else {
            const int rc = cdc::host::compute_amax(m, &amax);
            if (rc) return rc;
        }
        const double dt = cdc::host::step_dt(m, amax);
        if (dt_out) dt_out[runs[r]] = dt;
        cdc::host::plan_steps(m, 1, dt, plans[(size_t)r]);
        // tile offsets of each group from 0, one group after the other
        if (r == 0 || laws[(size_t)r] != laws[(size_t)r - 1]) {
            e->groups.push_back({r, r, laws[(size_t)r] / 4, laws[(size_t)r] % 4, 0});
            tile_begin.push_back(0);
        }
        cdc_ensemble::Group &G = e->groups.back();
        G.r1 = r + 1;
        G.ntiles += m->ntiles;
        tile_begin.push_back(G.ntiles);
        block_begin[(size_t)r + 1] = block_begin[(size_t)r] + m->nblk;
        for (int k = 0; k < e->rk; ++k) e->h_args[(size_t)k * nrun + r] = plans[(size_t)r].args[(size_t)k];
    }
    const int total_blocks = block_begin[(size_t)nrun];
    // marching members read the strip-edge columns of their committed state
    for (int r = 0; r < nrun; ++r) {
        const int rc = cdc::host::ensure_edges(e->members[(size_t)runs[r]], e->stream);
        if (rc) return rc;
    }

    // the graph is rebuilt when the active set or a law changed
    const bool same = e->graph_runs == runs && e->graph_laws == laws && e->exec;
    if (!same) {
        if (e->exec) (void)hipGraphExecDestroy(e->exec);
        if (e->graph) (void)hipGraphDestroy(e->graph);
        e->exec = nullptr;
        e->graph = nullptr;
        e->graph_runs.clear();
        e->graph_laws.clear();
        HIP_TRY(hipMemcpyAsync(e->d_tile_begin, tile_begin.data(), tile_begin.size() * sizeof(int32_t),
                               hipMemcpyHostToDevice, e->stream));
        HIP_TRY(hipMemcpyAsync(e->d_block_begin, block_begin.data(), (nrun + 1) * sizeof(int32_t),
                               hipMemcpyHostToDevice, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
        if (e->use_graph) {
            HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue_step(e, nrun, total_blocks);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(e->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (ec != hipSuccess) return set_msg(CDC_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
            e->graph = g;
            HIP_TRY(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
            e->graph_runs = runs;
            e->graph_laws = laws;
        }
    }
    if (e->use_graph) {
        HIP_TRY(hipGraphLaunch(e->exec, e->stream));
    } else {
        const int rc = enqueue_step(e, nrun, total_blocks);
        if (rc) return rc;
    }
    HIP_TRY(hipStreamSynchronize(e->stream));

    // the returned code is the lowest-indexed failing member's
    int first = CDC_OK, first_i = n;
    for (int r = 0; r < nrun; ++r) {
        const int i = runs[(size_t)r];
        cdc_solver *m = e->members[(size_t)i];
        const int flag = e->h_flags[r];
        if (flag == 0) {
            cdc::host::commit(m, plans[(size_t)r]);
            e->amax[(size_t)i] = e->h_amax[r];
            e->amax_of[(size_t)i] = m->version;
            e->steps += 1;
            e->cells += (double)m->ncell();
            continue;
        }
        e->amax_of[(size_t)i] = UINT64_MAX;
        const int code = cdc::host::resolve_error(m, flag, plans[(size_t)r], errs_out ? &errs_out[i] : nullptr);
        if (codes_out) codes_out[i] = code;
        if (i < first_i) {
            first = code;
            first_i = i;
        }
    }
    return first;
}

}  // extern "C"
