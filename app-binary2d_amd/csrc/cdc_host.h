// cdc_host.h -- host-side solver object and the pieces of one advance that the
// single-run path (cdc_solver.cpp) and the batched path (cdc_batch.cpp)
// share.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/cdc_solver.h"
#include "cdc_internal.h"

namespace cdc {
namespace host {

// Records the message returned by cdc_last_error_message and returns code.
int set_msg(int code, const std::string &m);

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return ::cdc::host::set_msg(CDC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int NBUF = 4;

// The state one RK stage reads or writes: buffer index plus the time and
// iteration carried with it.
struct Slot {
    int idx;
    double time;
    int64_t num, den;
};

// The stages of `fold` RK steps at one dt: launch arguments in launch order,
// the input time of every stage (the orbital state an error reports), and
// the slot that holds the result.
struct StepPlan {
    std::vector<StageArgs> args;
    std::vector<double> stage_times;
    Slot out;
};

}  // namespace host
}  // namespace cdc

struct cdc_solver {
    cdc_mesh_t mesh;
    cdc_hydro_t hydro;
    cdc_physics_t phys;
    int nv, nb, bs, N, nblk, ntiles, tpb, nsum, max_tiles;
    cdc::TileShape ts;
    hipStream_t stream = nullptr;
    int device = 0;

    double *d_init = nullptr;
    double *d_u[cdc::host::NBUF] = {};
    double *d_ist[cdc::host::NBUF] = {};
    double *d_oec[cdc::host::NBUF] = {};
    double *d_tq[cdc::host::NBUF] = {};   // [nblk][2] torque impulses (extension)
    // two buffers of [max_tiles][nsum + 1] per-tile sums, used by stage
    // parity, so that stage i + 1 writes while the finalize of stage i,
    // fused into the same launch, reads
    double *d_tile_sums = nullptr;
    size_t ts_stride = 0;
    const double *ts_last = nullptr;   // buffer of the last committed stage (max-speed column)
    cdc::TileErr *d_tile_err = nullptr;
    int32_t *d_perr = nullptr;
    int32_t *d_err_flag = nullptr;
    double *d_aux = nullptr;     // static per-cell planes (buffer rate, beta-cooling Omega)
    int naux = 1;
    double *d_stage = nullptr;   // AoS staging
    double *d_amax = nullptr;    // max-speed partials
    int namax = 0;

    int cur = 0;                 // committed buffer
    double time = 0.0;
    int64_t num = 0, den = 1;
    bool amax_valid = false;     // ts_last[:, nsum] hold the committed state's max speed
    uint64_t version = 0;        // bumped whenever the committed state changes

    // launch arguments of the stages of one advance: pinned staging and its
    // device copy, and the batch-of-one offsets {0, ntiles, 0, nblk}
    cdc::StageArgs *h_args = nullptr, *d_args = nullptr;
    size_t cap_args = 0;
    int32_t *h_begin = nullptr, *d_begin = nullptr;   // pinned staging and device copy
    int32_t *h_flag = nullptr;   // pinned read-back of the error flag
    // max signal speed of an advance's result, folded by its last finalize
    // into d_amax_adv and read back into pinned h_amax; amax_host holds it
    // for the committed state of version amax_host_version
    double *d_amax_adv = nullptr, *h_amax = nullptr;
    double amax_host = 0.0;
    uint64_t amax_host_version = UINT64_MAX;

    // the device work of one untimed advance as a replayable graph, per
    // stage count and tile shape (the arguments are re-read from the pinned
    // staging at every replay); replays start dependent launches back to back
    struct AdvanceGraph {
        size_t n;
        cdc::TileShape ts;
        hipGraphExec_t exec;
    };
    std::vector<AdvanceGraph> graphs;
    bool use_graph = true;
    // identity of this solver object (never reused) and a count of
    // drop_graphs calls: a graph captured elsewhere (the closure's corrected
    // advance) over this solver's buffers is valid while both are unchanged
    uint64_t uid = 0;
    uint64_t graph_epoch = 0;
    // strip-edge columns of each state buffer (block sizes that are multiples
    // of 64 only), and the committed-state version whose d_edge[cur] is
    // current
    double *d_edge[cdc::host::NBUF] = {};
    uint64_t edge_version = UINT64_MAX;
    // the column-marching kernel's static operands (StageArgs::stat; with
    // the edge arrays only), nstat pair planes
    double *d_stat = nullptr;
    int nstat = 0;

    bool timing = false;
    int64_t launches = 0;
    double total_ms = 0.0;
    std::vector<hipEvent_t> ev;

    size_t ncell() const { return (size_t)N * N; }
    size_t nstate() const { return ncell() * nv; }
};

namespace cdc {
namespace host {

// Stage (and, when finalize, per-block finalize) launches of args[0..n) for
// this one run (a batch of one), after one upload of the arguments; events
// ev[2i], ev[2i+1] bracket stage launch i when ev is given.  With fuse, the
// finalize of stage i runs as extra workgroups of stage i + 1's launch (one
// launch and one dependent-launch gap fewer per stage) and only the last
// stage's finalize is a launch of its own; a stage whose predecessor's
// finalize failed is then not reliably skipped, so an advance that raises
// its error flag is repeated unfused to name the reference's first error.
int launch_stages(cdc_solver *s, const StageArgs *args, size_t n, const cdc::TileShape &ts, hipEvent_t *ev,
                  bool finalize, bool fuse = false);
// Copies args[0..n) into the pinned staging, growing it (and dropping the
// cached graphs, which hold its old addresses) when needed.
int stage_args(cdc_solver *s, const StageArgs *args, size_t n);
// Enqueues the upload of the n staged stage arguments and the batch
// offsets, and with beta cooling the advance's relaxation factor (every stage
// of an advance has one dt).
int upload_stages(cdc_solver *s, size_t n);
// Enqueues the stage launches [i0, i1) of n uploaded stages: with fuse, each
// stage also finalizes the one before it (the last stage's finalize follows
// when i1 == n); without, each stage is followed by its own finalize.
int enqueue_stage_range(cdc_solver *s, size_t n, size_t i0, size_t i1, bool fuse);
// Drops the cached advance graphs.
void drop_graphs(cdc_solver *s);
// Clears the run's error flag.
int launch_set_int_checked(cdc_solver *s);
// Global max signal speed of the committed state (State::max_signal_speed).
int compute_amax(cdc_solver *s, double *out);
// dt = cfl * dx / amax, optionally raised to dt_min (src/scheme.rs:336-340).
double step_dt(const cdc_solver *s, double amax);
// Launch arguments of `fold` RK steps at dt from the committed state.
void plan_steps(const cdc_solver *s, int fold, double dt, StepPlan &plan);
// Makes the plan's result the committed state.
void commit(cdc_solver *s, const StepPlan &plan);
// Whether the stages launch the column-marching kernel (which reads and
// writes the strip-edge arrays).
bool marching(const cdc_solver *s);
// Before stages that march: enqueues, on stream st, the strip-edge columns
// of the committed state into d_edge[cur] unless they are current.
int ensure_edges(cdc_solver *s, hipStream_t st);
// After a failed advance (err_flag = flag): the error join_all would report.
int resolve_error(cdc_solver *s, int flag, const StepPlan &plan, cdc_error_t *err);

// Mean of r values x[0], x[stride], ... (r a power of two) by a pairwise tree.
double tree_mean(const double *x, int64_t stride, int64_t r);
// Per-block ist / oec of a re-gridded state: copied when num_blocks is kept,
// summed over merged blocks when a coarse block covers whole fine blocks,
// zero otherwise (either destination may be null).
void resample_block_sums(const cdc_mesh_t *src, const cdc_mesh_t *dst, int nv, const double *ist_src,
                         const double *oec_src, double *ist_dst, double *oec_dst);

}  // namespace host
}  // namespace cdc
