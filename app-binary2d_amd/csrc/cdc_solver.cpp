// cdc_solver.cpp -- host side of the C ABI (include/cdc_solver.h).
//
// Mirrors scheme::advance (src/scheme.rs:323-349) and the RK driver it runs:
// dt from the global max signal speed, then `fold` steps of RK1/2/3 where
// each stage is one fused HIP kernel (try_advance_rk, src/scheme.rs:88-173,
// with the RK combine of src/state.rs:183-219 folded in) followed by a small
// per-block finalize kernel.  Stage times and the Rational64 iteration are
// carried on the host exactly as State::weighted_average computes them, and
// the binary's orbital state for each stage is evaluated on the host.
//
// Transactional like the reference's clone-and-replace (src/main.rs:159-160):
// an advance works only in scratch buffers and commits by swapping an index,
// so a failing call leaves the committed state untouched.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/cdc_solver.h"
#include "cdc_host.h"
#include "cdc_internal.h"

using cdc::host::NBUF;
using cdc::host::set_msg;
using cdc::host::Slot;
using cdc::host::StepPlan;

namespace {

thread_local std::string g_msg;

constexpr double ORBITAL_PERIOD = 2.0 * 3.14159265358979323846;

int64_t gcd64(int64_t a, int64_t b) {
    if (a < 0) a = -a;
    if (b < 0) b = -b;
    while (b) { int64_t t = a % b; a = b; b = t; }
    return a ? a : 1;
}

void rat_norm(int64_t &n, int64_t &d) {
    int64_t g = gcd64(n, d);
    n /= g; d /= g;
    if (d < 0) { n = -n; d = -d; }
}

// Rational64: s1 * (1 - b) + s0 * b (src/state.rs:215)
void rat_wavg(int64_t n1, int64_t d1, int64_t n0, int64_t d0, int64_t bn, int64_t bd, int64_t &on,
              int64_t &od) {
    int64_t an = (bd - bn) * n1, ad = bd * d1;
    int64_t cn = bn * n0, cd = bd * d0;
    rat_norm(an, ad);
    rat_norm(cn, cd);
    on = an * cd + cn * ad;
    od = ad * cd;
    rat_norm(on, od);
}

double softening_length(const cdc_physics_t &p) {
    return p.has_softening_length ? p.softening_length : p.sink_radius;  // src/physics.rs:399-405
}

}  // namespace

int cdc::host::set_msg(int code, const std::string &m) {
    g_msg = m;
    return code;
}

namespace {

int validate(const cdc_mesh_t *m, const cdc_hydro_t *h, const cdc_physics_t *p) {
    if (!m || !h || !p) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null configuration pointer");
    if (m->num_blocks < 1) return set_msg(CDC_ERR_INVALID_ARGUMENT, "num_blocks must be >= 1");
    if (m->block_size < 8 || m->block_size % 8 != 0)
        return set_msg(CDC_ERR_INVALID_ARGUMENT, "block_size must be a positive multiple of 8");
    if ((int64_t)m->num_blocks * m->block_size > 65536)
        return set_msg(CDC_ERR_INVALID_ARGUMENT, "num_blocks * block_size must be <= 65536");
    if (!(m->domain_radius > 0.0)) return set_msg(CDC_ERR_INVALID_ARGUMENT, "domain_radius must be > 0");
    if (h->kind != CDC_ISOTHERMAL && h->kind != CDC_EULER)
        return set_msg(CDC_ERR_INVALID_ARGUMENT, "hydro.kind must be isothermal or euler");
    if (h->kind == CDC_EULER && !(h->gamma_law_index > 1.0))
        return set_msg(CDC_ERR_INVALID_ARGUMENT, "gamma_law_index must be > 1");
    if (h->kind == CDC_EULER && h->cooling_model == CDC_COOLING_BETA && !(h->cooling_beta > 0.0))
        return set_msg(CDC_ERR_INVALID_ARGUMENT, "cooling_beta must be > 0");
    if (p->rk_order < 1 || p->rk_order > 3) return set_msg(CDC_ERR_INVALID_ARGUMENT, "rk_order must be 1, 2 or 3");
    if (p->viscosity_model != CDC_VISC_CONSTANT_NU && p->viscosity_model != CDC_VISC_ALPHA)
        return set_msg(CDC_ERR_INVALID_ARGUMENT, "unknown viscosity model");
    if (p->sink_model != CDC_SINK_IMPARTIAL && p->sink_model != CDC_SINK_TORQUE_FREE)
        return set_msg(CDC_ERR_INVALID_ARGUMENT, "unknown sink model");
    return CDC_OK;
}

cdc::StageConsts make_consts(const cdc_solver *s, double time, double dt) {
    const cdc_physics_t &p = s->phys;
    const cdc_hydro_t &h = s->hydro;
    cdc::StageConsts k;
    memset(&k, 0, sizeof(k));
    const double L = 2.0 * s->mesh.domain_radius / (double)s->nb;
    k.dx = L / (double)s->bs;
    k.dy = k.dx;
    k.idx = 1.0 / k.dx;
    k.idy = 1.0 / k.dy;
    const double rs = softening_length(p);
    k.rs2 = rs * rs;
    k.inv_mach2 = h.kind == CDC_ISOTHERMAL ? 1.0 / (h.mach_number * h.mach_number) : 0.0;
    const double gm = h.kind == CDC_EULER ? h.gamma_law_index : 1.0;
    k.gm = gm;
    k.gm1 = gm - 1.0;
    k.inv_gm1 = h.kind == CDC_EULER ? 1.0 / (gm - 1.0) : 0.0;
    k.plm = p.plm;
    k.has_pdt = p.has_plm_density_trigger;
    k.pdt = p.has_plm_density_trigger ? p.plm_density_trigger : 0.0;
    k.viscosity_model = p.viscosity_model;
    k.visc = p.viscosity_model == CDC_VISC_ALPHA ? p.viscosity / sqrt(gm) : p.viscosity;
    k.lam = p.lambda_multiplier;
    k.eps_dx = 1e-12 * k.dx;
    double el[4];
    cdc::orbital_elements(p, el);
    k.sink_s2 = p.sink_radius * p.sink_radius;
    k.inv_sink_s2 = 1.0 / k.sink_s2;
    k.sink_rate = p.sink_rate;
    k.fake_threshold = p.fake_mass_threshold;
    k.fake_rate = p.fake_mass_rate;
    k.cooling_coefficient = h.cooling_coefficient;
    k.cooling_inv_beta = h.cooling_model == CDC_COOLING_BETA ? 1.0 / h.cooling_beta : 0.0;
    k.cooling_f = h.cooling_slow_start ? exp(-ORBITAL_PERIOD / (time + dt)) : 1.0;
    k.has_mach_ceiling = p.has_mach_ceiling;
    k.mach_ceiling_coef = p.has_mach_ceiling ? 2.0 / gm / (gm - 1.0) / (p.mach_ceiling * p.mach_ceiling) : 0.0;
    k.has_density_floor = h.has_density_floor;
    k.density_floor = h.density_floor;
    k.density_floor_pressure = (h.kind == CDC_EULER && h.has_density_floor) ? pow(h.density_floor, gm) : 0.0;
    k.has_pressure_floor = h.kind == CDC_EULER ? h.has_pressure_floor : 0;
    k.pressure_floor = h.pressure_floor;
    k.density_floor_cmp = h.has_density_floor ? h.density_floor : -HUGE_VAL;
    k.pressure_floor_cmp = k.has_pressure_floor ? h.pressure_floor : -HUGE_VAL;
    k.density_index = h.density_index;
    k.cooling_slow_start = h.cooling_slow_start;
    k.cooling_model = h.cooling_model;
    k.sink_model = p.sink_model;
    k.one_body = p.one_body;
    return k;
}

void free_all(cdc_solver *s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    cdc::host::drop_graphs(s);
    (void)hipFree(s->d_init);
    for (int i = 0; i < NBUF; ++i) {
        (void)hipFree(s->d_u[i]);
        (void)hipFree(s->d_ist[i]);
        (void)hipFree(s->d_oec[i]);
        (void)hipFree(s->d_tq[i]);
        (void)hipFree(s->d_edge[i]);
    }
    (void)hipFree(s->d_tile_sums);
    (void)hipFree(s->d_tile_err);
    (void)hipFree(s->d_perr);
    (void)hipFree(s->d_err_flag);
    (void)hipFree(s->d_aux);
    (void)hipFree(s->d_stat);
    (void)hipFree(s->d_stage);
    (void)hipFree(s->d_amax);
    (void)hipFree(s->d_args);
    (void)hipFree(s->d_begin);
    if (s->h_args) (void)hipHostFree(s->h_args);
    if (s->h_begin) (void)hipHostFree(s->h_begin);
    if (s->h_flag) (void)hipHostFree(s->h_flag);
    if (s->h_amax) (void)hipHostFree(s->h_amax);
    (void)hipFree(s->d_amax_adv);
    for (auto e : s->ev) (void)hipEventDestroy(e);
    if (s->stream) (void)hipStreamDestroy(s->stream);
}

// Host-side cell centre, as Mesh::cell_centers (src/mesh.rs:67-72)
void cell_centre(const cdc_solver *s, int b, int cell, double *x, double *y) {
    const double R = s->mesh.domain_radius;
    const double L = 2.0 * R / (double)s->nb;
    const int bi = b / s->nb, bj = b % s->nb;
    const int i = cell / s->bs, j = cell % s->bs;
    auto axis = [&](int blk, int k) {
        const double a = -R + (double)blk * L;
        const double e = a + L;
        const double step = (e - a) / (double)s->bs;
        const double v0 = a + step * (double)k;
        const double v1 = a + step * (double)(k + 1);
        return (v0 + v1) / 2.0;
    };
    *x = axis(bi, i);
    *y = axis(bj, j);
}

int upload_conserved(cdc_solver *s, const double *aos, double *soa) {
    HIP_TRY(hipMemcpyAsync(s->d_stage, aos, s->nstate() * sizeof(double), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(cdc::launch_aos_to_soa(s->nv, s->nb, s->bs, s->d_stage, soa, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return CDC_OK;
}

// Static per-cell planes for the current physics (buffer rate, Omega)
int compute_aux(cdc_solver *s) {
    double el[4];
    cdc::orbital_elements(s->phys, el);
    const double R = s->mesh.domain_radius;
    cdc::AuxArgs a;
    memset(&a, 0, sizeof(a));
    a.aux = s->d_aux;
    a.N = s->N;
    a.nb = s->nb;
    a.bs = s->bs;
    a.naux = s->naux;
    a.R = R;
    a.L = 2.0 * R / (double)s->nb;
    a.buffer_rate = s->phys.buffer_rate;
    a.buffer_scale = s->phys.buffer_scale;
    a.omega_outer = sqrt(el[1] / (R * R * R));
    a.rs2 = softening_length(s->phys) * softening_length(s->phys);
    a.mtot = el[1];
    HIP_TRY(cdc::launch_aux(a, s->stream));
    if (s->d_stat)
        HIP_TRY(cdc::launch_stat_build(s->nv, s->nstat, s->d_init, s->d_aux, s->d_stat, (int64_t)s->ncell(),
                                       s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return CDC_OK;
}

}  // namespace

namespace cdc {
namespace host {

int launch_set_int_checked(cdc_solver *s) {
    HIP_TRY(cdc::launch_set_int(s->d_err_flag, 0, s->stream));
    return CDC_OK;
}

void drop_graphs(cdc_solver *s) {
    for (auto &g : s->graphs) (void)hipGraphExecDestroy(g.exec);
    s->graphs.clear();
    s->graph_epoch += 1;
}

int stage_args(cdc_solver *s, const StageArgs *args, size_t n) {
    if (n > s->cap_args) {
        HIP_TRY(hipStreamSynchronize(s->stream));
        drop_graphs(s);
        if (s->h_args) (void)hipHostFree(s->h_args);
        if (s->d_args) (void)hipFree(s->d_args);
        s->h_args = nullptr;
        s->d_args = nullptr;
        s->cap_args = 0;
        const size_t cap = std::max<size_t>(n, 32);
        HIP_TRY(hipHostMalloc((void **)&s->h_args, cap * sizeof(StageArgs), hipHostMallocDefault));
        HIP_TRY(hipMalloc((void **)&s->d_args, cap * sizeof(StageArgs)));
        s->cap_args = cap;
    }
    // the staging buffer is free: every earlier call synchronised its stream
    memcpy(s->h_args, args, n * sizeof(StageArgs));
    s->h_begin[0] = 0;
    s->h_begin[1] = args[0].ntiles;
    s->h_begin[2] = 0;
    s->h_begin[3] = s->nblk;
    return CDC_OK;
}

// The stage launches of n staged stages: argument upload, then per stage the
// stage kernel (fused: also finalizing the stage before) and, unfused, its
// finalize; fused, the last stage's finalize at the end.  Timing events are
// bound to each stage dispatch.
int upload_stages(cdc_solver *s, size_t n) {
    HIP_TRY(hipMemcpyAsync(s->d_args, s->h_args, n * sizeof(StageArgs), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_begin, s->h_begin, 4 * sizeof(int32_t), hipMemcpyHostToDevice, s->stream));
    // every stage of an advance has the first one's dt; the factor goes where
    // the kernel that runs the stages reads it
    if (s->h_args[0].cool_factor)
        HIP_TRY(cdc::launch_cooling_factor(s->d_args, s->d_aux, marching(s) ? s->d_stat : nullptr,
                                           (int64_t)s->ncell(), s->stream));
    return CDC_OK;
}

// the run's viscosity law (its physics may change between calls; every
// change drops the cached graphs)
static int stage_viscosity(const cdc_solver *s) {
    return s->phys.viscosity_model == CDC_VISC_ALPHA ? cdc::kViscAlpha : cdc::kViscConst;
}

// the stage's cooling law when it is fixed for the launch
static int stage_cooling(const cdc_solver *s, const StageArgs &h) {
    return s->nv == 3 ? cdc::kCoolAny
           : s->hydro.cooling_model == CDC_COOLING_BETA ? (h.cool_factor ? cdc::kCoolBetaFactor : cdc::kCoolBetaExp)
                                                        : cdc::kCoolT4;
}

int enqueue_stage_range(cdc_solver *s, size_t n, size_t i0, size_t i1, bool fuse) {
    const int ntiles = s->h_args[0].ntiles;
    for (size_t i = i0; i < i1; ++i) {
        const StageArgs *fin = fuse && i > 0 ? s->d_args + i - 1 : nullptr;
        const StageArgs &h = s->h_args[i];
        HIP_TRY(cdc::launch_stage_batch(s->nv, s->ts, s->d_args + i, s->d_begin, 1, ntiles, fin, s->d_begin + 2,
                                        s->nblk, stage_cooling(s, h), stage_viscosity(s), h.read_base, nullptr,
                                        nullptr, s->stream));
        if (!fuse) HIP_TRY(cdc::launch_finalize_batch(s->nv, s->d_args + i, s->d_begin + 2, 1, s->nblk, s->stream));
    }
    if (fuse && i1 == n && n > 0)
        HIP_TRY(cdc::launch_finalize_batch(s->nv, s->d_args + n - 1, s->d_begin + 2, 1, s->nblk, s->stream));
    return CDC_OK;
}

static int enqueue_stages(cdc_solver *s, size_t n, int ntiles, const cdc::TileShape &ts, hipEvent_t *ev,
                          bool finalize, bool fuse) {
    int rc = upload_stages(s, n);
    if (rc) return rc;
    fuse = fuse && finalize;
    for (size_t i = 0; i < n; ++i) {
        // fused: this launch also finalizes stage i - 1 (its tile sums are in
        // the other parity buffer, its ist / oec inputs are complete)
        const StageArgs *fin = fuse && i > 0 ? s->d_args + i - 1 : nullptr;
        const StageArgs &h = s->h_args[i];
        HIP_TRY(cdc::launch_stage_batch(s->nv, ts, s->d_args + i, s->d_begin, 1, ntiles, fin, s->d_begin + 2,
                                        s->nblk, stage_cooling(s, h), stage_viscosity(s), h.read_base, ev ? ev[2 * i] : nullptr,
                                        ev ? ev[2 * i + 1] : nullptr, s->stream));
        if (finalize && !fuse)
            HIP_TRY(cdc::launch_finalize_batch(s->nv, s->d_args + i, s->d_begin + 2, 1, s->nblk, s->stream));
    }
    if (fuse && n > 0)
        HIP_TRY(cdc::launch_finalize_batch(s->nv, s->d_args + n - 1, s->d_begin + 2, 1, s->nblk, s->stream));
    return CDC_OK;
}

int launch_stages(cdc_solver *s, const StageArgs *args, size_t n, const cdc::TileShape &ts, hipEvent_t *ev,
                  bool finalize, bool fuse) {
    int rc = stage_args(s, args, n);
    if (rc) return rc;
    return enqueue_stages(s, n, args[0].ntiles, ts, ev, finalize, fuse);
}

// One advance's device work, fused, from the staged arguments: flag reset,
// stages, the flag read into s->h_flag.  Replayed from a cached graph when
// capture works on this device, else launched directly.  With timing events
// it is launched directly, the events bound to each stage dispatch: events
// recorded as nodes of a captured graph cannot be timed on ROCm.
static int enqueue_advance(cdc_solver *s, size_t n, int ntiles, hipEvent_t *ev) {
    const bool timed = ev != nullptr;
    if (s->use_graph && !timed) {
        for (const auto &g : s->graphs)
            if (g.n == n && g.ts.tx == s->ts.tx && g.ts.ty == s->ts.ty && g.ts.nt == s->ts.nt) {
                HIP_TRY(hipGraphLaunch(g.exec, s->stream));
                return CDC_OK;
            }
        int rc = CDC_OK;
        if (hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            s->use_graph = false;
            return enqueue_advance(s, n, ntiles, ev);
        }
        if (cdc::launch_reset_advance(s->d_err_flag, s->d_amax_adv, s->stream) != hipSuccess)
            rc = set_msg(CDC_ERR_HIP, "reset");
        if (!rc) rc = enqueue_stages(s, n, ntiles, s->ts, nullptr, true, true);
        if (!rc && (hipMemcpyAsync(s->h_flag, s->d_err_flag, sizeof(int32_t), hipMemcpyDeviceToHost, s->stream) !=
                        hipSuccess ||
                    hipMemcpyAsync(s->h_amax, s->d_amax_adv, sizeof(double), hipMemcpyDeviceToHost, s->stream) !=
                        hipSuccess))
            rc = set_msg(CDC_ERR_HIP, "read-back");
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(s->stream, &graph);
        hipGraphExec_t exec = nullptr;
        if (!rc && ec == hipSuccess && graph &&
            hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess) {
            (void)hipGraphDestroy(graph);
            if (s->graphs.size() >= 4) {
                (void)hipGraphExecDestroy(s->graphs.front().exec);
                s->graphs.erase(s->graphs.begin());
            }
            s->graphs.push_back({n, s->ts, exec});
            HIP_TRY(hipGraphLaunch(exec, s->stream));
            return CDC_OK;
        }
        // capture is not available here: launch directly from now on
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        s->use_graph = false;
    }
    HIP_TRY(cdc::launch_reset_advance(s->d_err_flag, s->d_amax_adv, s->stream));
    int rc = enqueue_stages(s, n, ntiles, s->ts, ev, true, true);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(s->h_flag, s->d_err_flag, sizeof(int32_t), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(s->h_amax, s->d_amax_adv, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    return CDC_OK;
}

int compute_amax(cdc_solver *s, double *out) {
    if (s->amax_host_version == s->version) {
        // left by the advance that produced the committed state
        *out = s->amax_host;
        return CDC_OK;
    }
    std::vector<double> h;
    if (s->amax_valid) {
        // max signal speed of the committed state, left by the last stage
        const int nr = s->nsum + 1;
        double a = 0.0;
        HIP_TRY(cdc::launch_column_max(s->ts_last + s->nsum, s->ntiles, nr, s->d_amax, s->stream));
        HIP_TRY(hipMemcpyAsync(&a, s->d_amax, sizeof(double), hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        *out = a;
        return CDC_OK;
    }
    HIP_TRY(cdc::launch_max_speed(s->nv, s->d_u[s->cur], (int64_t)s->ncell(),
                                  s->hydro.kind == CDC_EULER ? s->hydro.gamma_law_index : 1.0, s->d_amax, s->namax,
                                  s->stream));
    h.resize(s->namax);
    HIP_TRY(hipMemcpyAsync(h.data(), s->d_amax, h.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    double a = 0.0;
    for (double v : h) a = std::max(v, a);
    *out = a;
    return CDC_OK;
}

double step_dt(const cdc_solver *s, double amax) {
    const cdc_physics_t &ph = s->phys;
    const double L = 2.0 * s->mesh.domain_radius / (double)s->nb;
    const double dx = L / (double)s->bs;
    double dt = ph.cfl * dx / amax;
    if (ph.has_dt_min) dt = std::max(dt, ph.dt_min);
    return dt;
}

void plan_steps(const cdc_solver *s, int fold, double dt, StepPlan &plan) {
    const cdc_physics_t &ph = s->phys;
    const double L = 2.0 * s->mesh.domain_radius / (double)s->nb;
    double el[4];
    cdc::orbital_elements(ph, el);
    const int rk = ph.rk_order;
    plan.args.clear();
    plan.stage_times.clear();
    plan.args.reserve((size_t)fold * rk);
    plan.stage_times.reserve((size_t)fold * rk);

    auto pick = [&](int a, int b, int c) {
        for (int i = 0; i < NBUF; ++i)
            if (i != a && i != b && i != c) return i;
        return -1;
    };

    // one stage: out = L(in) combined with base when read_base
    auto stage = [&](const Slot &in, const Slot &base, bool read_base, int64_t bn, int64_t bd, bool last,
                     Slot &out) {
        const int sid = (int)plan.args.size() + 1;
        plan.stage_times.push_back(in.time);
        out.idx = pick(in.idx, base.idx, s->cur);
        const double bf = read_base ? (double)bn / (double)bd : 0.0;
        const double c1 = -bf + 1.0;

        cdc::StageArgs a;
        memset(&a, 0, sizeof(a));
        a.uc = s->d_u[in.idx];
        a.u0 = s->d_init;
        a.stat = s->d_stat;
        a.ub = s->d_u[base.idx];
        a.uo = s->d_u[out.idx];
        a.aux = s->d_aux;
        a.tile_sums = s->d_tile_sums + (size_t)((sid - 1) & 1) * s->ts_stride;
        a.tile_err = s->d_tile_err;
        a.err_flag = s->d_err_flag;
        a.ist_c = s->d_ist[in.idx];
        a.ist_b = s->d_ist[base.idx];
        a.oec_c = s->d_oec[in.idx];
        a.oec_b = s->d_oec[base.idx];
        a.ist_o = s->d_ist[out.idx];
        a.oec_o = s->d_oec[out.idx];
        a.tq_c = s->d_tq[in.idx];
        a.tq_b = s->d_tq[base.idx];
        a.tq_o = s->d_tq[out.idx];
        a.perr = s->d_perr;
        a.ec = s->d_edge[in.idx];
        a.eo = s->d_edge[out.idx];
        a.reuse_prim = s->hydro.has_density_floor ? 0 : 1;
        a.N = s->N;
        a.nb = s->nb;
        a.bs = s->bs;
        a.ntiles = s->ntiles;
        a.tpb = s->tpb;
        a.stage_id = sid;
        a.read_base = read_base ? 1 : 0;
        a.want_amax = last ? 1 : 0;
        a.R = s->mesh.domain_radius;
        a.L = L;
        a.time = in.time;
        a.dt = dt;
        a.c1 = c1;
        a.c0 = bf;
        a.orb = cdc::orbital_state(el, in.time);
        a.k = make_consts(s, in.time, dt);
        a.area = a.k.dx * a.k.dy;
        for (int i = 0; i < 4; ++i) a.el[i] = el[i];
        plan.args.push_back(a);

        // time and iteration of the stage result, then of the combine
        double t1 = in.time + dt;
        int64_t n1 = in.num + in.den, d1 = in.den;
        rat_norm(n1, d1);
        if (read_base) {
            out.time = t1 * c1 + base.time * bf;
            rat_wavg(n1, d1, base.num, base.den, bn, bd, out.num, out.den);
        } else {
            out.time = t1;
            out.num = n1;
            out.den = d1;
        }
    };

    Slot cur{s->cur, s->time, s->num, s->den};
    for (int f = 0; f < fold; ++f) {
        const bool lastf = f == fold - 1;
        Slot s1, s2, s3;
        if (rk == 1) {
            stage(cur, cur, false, 0, 1, lastf, s1);
            cur = s1;
        } else if (rk == 2) {
            stage(cur, cur, false, 0, 1, false, s1);
            stage(s1, cur, true, 1, 2, lastf, s2);
            cur = s2;
        } else {
            stage(cur, cur, false, 0, 1, false, s1);
            stage(s1, cur, true, 3, 4, false, s2);
            stage(s2, cur, true, 1, 3, lastf, s3);
            cur = s3;
        }
    }
    plan.out = cur;
}

void commit(cdc_solver *s, const StepPlan &plan) {
    const Slot &out = plan.out;
    s->ts_last = plan.args.back().tile_sums;
    s->cur = out.idx;
    s->time = out.time;
    s->num = out.num;
    s->den = out.den;
    s->amax_valid = true;
    s->version += 1;
    // a marching stage wrote the result's strip-edge columns
    if (marching(s)) s->edge_version = s->version;
}

bool marching(const cdc_solver *s) { return cdc::march_shape(s->ts) && s->d_edge[0] != nullptr; }

int ensure_edges(cdc_solver *s, hipStream_t st) {
    if (!marching(s) || s->edge_version == s->version) return CDC_OK;
    HIP_TRY(cdc::launch_edge_build(s->nv, s->N, s->d_u[s->cur], s->d_edge[s->cur], st));
    s->edge_version = s->version;
    return CDC_OK;
}

int resolve_error(cdc_solver *s, int flag, const StepPlan &plan, cdc_error_t *err) {
    // Resolve the error as join_all over blocks would (src/scheme.rs:161-165):
    // block b fails with the first failing 3x3 neighbour's P-phase error
    // (try_join_3by3 order), else with its own perturbation error.
    s->amax_valid = false;
    if (flag < 1 || (size_t)flag > plan.stage_times.size())
        return set_msg(CDC_ERR_HIP, "error flag names no stage of the call");
    std::vector<cdc::TileErr> te((size_t)s->ntiles);
    std::vector<int32_t> pe((size_t)s->nblk);
    HIP_TRY(hipMemcpy(te.data(), s->d_tile_err, te.size() * sizeof(cdc::TileErr), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(pe.data(), s->d_perr, pe.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<cdc::TileErr> berr((size_t)s->nblk, cdc::TileErr{0, 0, 0.0});
    for (int b = 0; b < s->nblk; ++b)
        for (int t = 0; t < s->tpb; ++t) {
            const cdc::TileErr &e = te[(size_t)b * s->tpb + t];
            if (e.code && (!berr[b].code || e.cell < berr[b].cell)) berr[b] = e;
        }
    const double stage_time = plan.stage_times[(size_t)flag - 1];
    double el[4];
    cdc::orbital_elements(s->phys, el);
    cdc::Orbit orb = cdc::orbital_state(el, stage_time);
    for (int b = 0; b < s->nblk; ++b) {
        const int bi = b / s->nb, bj = b % s->nb;
        int code = 0, nbk = -1;
        for (int r = 0; r < 3 && !code; ++r)
            for (int c = 0; c < 3 && !code; ++c) {
                const int n = ((bi + s->nb + r - 1) % s->nb) * s->nb + (bj + s->nb + c - 1) % s->nb;
                if (berr[n].code) { code = berr[n].code; nbk = n; }
            }
        if (!code && pe[b]) code = pe[b];
        if (!code) continue;
        if (err) {
            err->code = code;
            err->time = stage_time;
            err->has_binary = 1;
            err->binary[0] = {orb.m1, orb.x1, orb.y1, orb.vx1, orb.vy1};
            err->binary[1] = {orb.m2, orb.x2, orb.y2, orb.vx2, orb.vy2};
            if (nbk >= 0) {
                err->value = berr[nbk].value;
                cell_centre(s, nbk, berr[nbk].cell, &err->x, &err->y);
            }
            snprintf(err->message, sizeof(err->message), "%s %.4e at position (%.4f %.4f)", cdc_error_string(code),
                     err->value, err->x, err->y);
        }
        return set_msg(code, cdc_error_string(code));
    }
    return set_msg(CDC_ERR_HIP, "error flag set without a failing block");
}

}  // namespace host
}  // namespace cdc

using cdc::host::compute_amax;

extern "C" {

int32_t cdc_abi_version(void) { return CDC_ABI_VERSION; }

int32_t cdc_struct_size(int32_t which) {
    switch (which) {
        case 0: return (int32_t)sizeof(cdc_mesh_t);
        case 1: return (int32_t)sizeof(cdc_hydro_t);
        case 2: return (int32_t)sizeof(cdc_physics_t);
        case 3: return (int32_t)sizeof(cdc_error_t);
        case 4: return (int32_t)sizeof(cdc_cavity_model_t);
        case 5: return (int32_t)sizeof(cdc_closure_config_t);
        default: return -1;
    }
}

int32_t cdc_math_probe(const double *x, double *out, int32_t n) {
    if (!x || !out || n < 0) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null pointer or negative count");
    if (n == 0) return CDC_OK;
    double *dx = nullptr, *dout = nullptr;
    if (hipMalloc((void **)&dx, (size_t)n * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&dout, (size_t)n * 3 * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(dx);
        return set_msg(CDC_ERR_OUT_OF_MEMORY, "math probe buffers");
    }
    hipError_t e = hipMemcpy(dx, x, (size_t)n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = cdc::launch_math_probe(dx, dout, n, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out, dout, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(dx);
    (void)hipFree(dout);
    if (e != hipSuccess) return set_msg(CDC_ERR_HIP, std::string("math probe: ") + hipGetErrorString(e));
    return CDC_OK;
}

int32_t cdc_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

const char *cdc_error_string(int32_t code) {
    switch (code) {
        case CDC_OK: return "ok";
        case CDC_ERR_NEGATIVE_DENSITY: return "negative surface density";
        case CDC_ERR_NEGATIVE_PRESSURE: return "negative gas pressure";
        case CDC_ERR_UNBOUND_ORBIT: return "unbound orbital state";
        case CDC_ERR_INVALID_ARGUMENT: return "invalid argument";
        case CDC_ERR_HIP: return "HIP runtime error";
        case CDC_ERR_OUT_OF_MEMORY: return "out of device memory";
        case CDC_ERR_NO_DEVICE: return "no HIP device";
        default: return "unknown error";
    }
}

const char *cdc_last_error_message(void) { return g_msg.c_str(); }

int32_t cdc_orbital_state_from_time(const cdc_physics_t *physics, double time, cdc_point_mass_t out[2]) {
    if (!physics || !out) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null pointer");
    double el[4];
    cdc::orbital_elements(*physics, el);
    cdc::Orbit o = cdc::orbital_state(el, time);
    out[0] = {o.m1, o.x1, o.y1, o.vx1, o.vy1};
    out[1] = {o.m2, o.x2, o.y2, o.vx2, o.vy2};
    return CDC_OK;
}

int32_t cdc_solver_create(const cdc_mesh_t *mesh, const cdc_hydro_t *hydro, const cdc_physics_t *physics,
                          const double *initial_conserved, cdc_solver_t **out) {
    if (!out) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    int rc = validate(mesh, hydro, physics);
    if (rc) return rc;
    if (!initial_conserved) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null initial_conserved");
    if (cdc_device_count() < 1) return set_msg(CDC_ERR_NO_DEVICE, "no HIP device visible");

    cdc_solver *s = new cdc_solver();
    static std::atomic<uint64_t> next_uid{1};
    s->uid = next_uid.fetch_add(1);
    s->mesh = *mesh;
    s->hydro = *hydro;
    s->phys = *physics;
    s->nv = hydro->kind == CDC_EULER ? 4 : 3;
    s->nb = mesh->num_blocks;
    s->bs = mesh->block_size;
    s->N = s->nb * s->bs;
    s->nblk = s->nb * s->nb;
    s->ts = cdc::tile_shape_for(s->bs, s->nv);
    s->tpb = (s->bs / s->ts.tx) * (s->bs / s->ts.ty);
    s->ntiles = s->nblk * s->tpb;
    s->max_tiles = s->ntiles;
    {
        const cdc::TileShape *shapes;
        const int n = cdc::tile_shapes(&shapes);
        for (int i = 0; i < n; ++i)
            if (s->bs % shapes[i].tx == 0 && s->bs % shapes[i].ty == 0)
                s->max_tiles = std::max(s->max_tiles, s->nblk * (s->bs / shapes[i].tx) * (s->bs / shapes[i].ty));
    }
    s->nsum = cdc::nsum_for(s->nv);
    (void)hipGetDevice(&s->device);

    auto fail = [&](int code, const std::string &m) {
        free_all(s);
        delete s;
        return set_msg(code, m);
    };
#define ALLOC(ptr, bytes)                                                                   \
    do {                                                                                    \
        if (hipMalloc((void **)&(ptr), (bytes)) != hipSuccess) {                            \
            (void)hipGetLastError();                                                        \
            return fail(CDC_ERR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(bytes) + " bytes"); \
        }                                                                                   \
    } while (0)

    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(CDC_ERR_HIP, "hipStreamCreate failed");
    const size_t sb = s->nstate() * sizeof(double);
    ALLOC(s->d_init, sb);
    ALLOC(s->d_stage, sb);
    for (int i = 0; i < NBUF; ++i) {
        ALLOC(s->d_u[i], sb);
        ALLOC(s->d_ist[i], (size_t)s->nblk * 7 * s->nv * sizeof(double));
        ALLOC(s->d_oec[i], (size_t)s->nblk * 7 * 4 * sizeof(double));
        ALLOC(s->d_tq[i], (size_t)s->nblk * 2 * sizeof(double));
        if (s->bs % 64 == 0 && cdc::march_fits(s->nv, s->N))
            ALLOC(s->d_edge[i], (size_t)s->nv * s->N * (s->N / 64) * 4 * sizeof(double));
    }
    s->ts_stride = (size_t)s->max_tiles * (s->nsum + 1);
    ALLOC(s->d_tile_sums, 2 * s->ts_stride * sizeof(double));
    s->ts_last = s->d_tile_sums;
    ALLOC(s->d_tile_err, (size_t)s->max_tiles * sizeof(cdc::TileErr));
    ALLOC(s->d_perr, (size_t)s->nblk * sizeof(int32_t));
    ALLOC(s->d_err_flag, sizeof(int32_t));
    // planes: buffer rate; with beta cooling also Omega and the per-advance
    // relaxation factor exp(-dt Omega / beta)
    s->naux = (hydro->kind == CDC_EULER && hydro->cooling_model == CDC_COOLING_BETA) ? 3 : 1;
    ALLOC(s->d_aux, (size_t)s->naux * s->ncell() * sizeof(double));
    if (s->d_edge[0]) {
        s->nstat = s->naux > 2 ? 3 : 2;
        ALLOC(s->d_stat, (size_t)s->nstat * 2 * s->ncell() * sizeof(double));
    }
    s->namax = 1024;
    ALLOC(s->d_amax, (size_t)s->namax * sizeof(double));
    ALLOC(s->d_begin, 4 * sizeof(int32_t));
    ALLOC(s->d_amax_adv, sizeof(double));
    if (hipHostMalloc((void **)&s->h_begin, 4 * sizeof(int32_t), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&s->h_flag, sizeof(int32_t), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&s->h_amax, sizeof(double), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return fail(CDC_ERR_OUT_OF_MEMORY, "pinned staging");
    }
#undef ALLOC

    rc = upload_conserved(s, initial_conserved, s->d_init);
    if (rc) return fail(rc, g_msg);
    if (hipMemcpyAsync(s->d_u[0], s->d_init, sb, hipMemcpyDeviceToDevice, s->stream) != hipSuccess ||
        hipMemsetAsync(s->d_ist[0], 0, (size_t)s->nblk * 7 * s->nv * sizeof(double), s->stream) != hipSuccess ||
        hipMemsetAsync(s->d_oec[0], 0, (size_t)s->nblk * 7 * 4 * sizeof(double), s->stream) != hipSuccess ||
        hipMemsetAsync(s->d_tq[0], 0, (size_t)s->nblk * 2 * sizeof(double), s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return fail(CDC_ERR_HIP, "initial state upload failed");
    rc = compute_aux(s);
    if (rc) return fail(rc, g_msg);
    s->cur = 0;
    s->time = 0.0;
    s->num = 0;
    s->den = 1;
    *out = s;
    return CDC_OK;
}

void cdc_solver_destroy(cdc_solver_t *s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    free_all(s);
    delete s;
}

int32_t cdc_solver_set_physics(cdc_solver_t *s, const cdc_physics_t *physics) {
    if (!s) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null solver");
    int rc = validate(&s->mesh, &s->hydro, physics);
    if (rc) return rc;
    if (physics->viscosity_model != s->phys.viscosity_model) {
        // the captured advances launch a kernel compiled for the old law
        HIP_TRY(hipStreamSynchronize(s->stream));
        cdc::host::drop_graphs(s);
    }
    s->phys = *physics;
    return compute_aux(s);
}

int32_t cdc_solver_set_state(cdc_solver_t *s, double time, int64_t num, int64_t den, const double *conserved,
                             const double *ist, const double *oec) {
    if (!s || !conserved) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null pointer");
    if (den == 0) return set_msg(CDC_ERR_INVALID_ARGUMENT, "iteration denominator is zero");
    int rc = upload_conserved(s, conserved, s->d_u[s->cur]);
    if (rc) return rc;
    const size_t ni = (size_t)s->nblk * 7 * s->nv * sizeof(double);
    const size_t ne = (size_t)s->nblk * 7 * 4 * sizeof(double);
    if (ist) HIP_TRY(hipMemcpyAsync(s->d_ist[s->cur], ist, ni, hipMemcpyHostToDevice, s->stream));
    else HIP_TRY(hipMemsetAsync(s->d_ist[s->cur], 0, ni, s->stream));
    if (oec) HIP_TRY(hipMemcpyAsync(s->d_oec[s->cur], oec, ne, hipMemcpyHostToDevice, s->stream));
    else HIP_TRY(hipMemsetAsync(s->d_oec[s->cur], 0, ne, s->stream));
    HIP_TRY(hipMemsetAsync(s->d_tq[s->cur], 0, (size_t)s->nblk * 2 * sizeof(double), s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->time = time;
    s->num = num;
    s->den = den;
    rat_norm(s->num, s->den);
    s->amax_valid = false;
    s->version += 1;
    return CDC_OK;
}

int32_t cdc_solver_get_state(cdc_solver_t *s, double *time, int64_t *num, int64_t *den, double *conserved,
                             double *ist, double *oec) {
    if (!s) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null solver");
    if (time) *time = s->time;
    if (num) *num = s->num;
    if (den) *den = s->den;
    if (conserved) {
        HIP_TRY(cdc::launch_soa_to_aos(s->nv, s->nb, s->bs, s->d_u[s->cur], s->d_stage, s->stream));
        HIP_TRY(hipMemcpyAsync(conserved, s->d_stage, s->nstate() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    }
    if (ist)
        HIP_TRY(hipMemcpyAsync(ist, s->d_ist[s->cur], (size_t)s->nblk * 7 * s->nv * sizeof(double),
                               hipMemcpyDeviceToHost, s->stream));
    if (oec)
        HIP_TRY(hipMemcpyAsync(oec, s->d_oec[s->cur], (size_t)s->nblk * 7 * 4 * sizeof(double),
                               hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return CDC_OK;
}

int32_t cdc_solver_set_kernel(cdc_solver_t *s, int32_t tx, int32_t ty, int32_t threads) {
    if (!s) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null solver");
    cdc::TileShape t = cdc::tile_shape_for(s->bs, s->nv);
    if (tx != 0 || ty != 0) {
        const cdc::TileShape *shapes;
        const int n = cdc::tile_shapes(&shapes);
        bool found = false;
        for (int i = 0; i < n && !found; ++i)
            if (shapes[i].tx == tx && shapes[i].ty == ty && (threads == 0 || shapes[i].nt == threads)) {
                t = shapes[i];
                found = true;
            }
        if (!found) return set_msg(CDC_ERR_INVALID_ARGUMENT, "unsupported tile shape");
        if (s->bs % tx != 0 || s->bs % ty != 0)
            return set_msg(CDC_ERR_INVALID_ARGUMENT, "tile shape must divide block_size");
        if (cdc::march_shape(t) && !cdc::march_fits(s->nv, s->N))
            return set_msg(CDC_ERR_INVALID_ARGUMENT,
                           "grid too large for the column-marching strips (arrays of 2^32 bytes or more): "
                           "use a tile shape");
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->ts = t;
    s->tpb = (s->bs / t.tx) * (s->bs / t.ty);
    s->ntiles = s->nblk * s->tpb;
    s->amax_valid = false;   // the per-tile max-speed column was written with the old tiling
    return CDC_OK;
}

int32_t cdc_solver_set_tile(cdc_solver_t *s, int32_t tx, int32_t ty) { return cdc_solver_set_kernel(s, tx, ty, 0); }

int32_t cdc_solver_get_kernel(cdc_solver_t *s, int32_t *tx, int32_t *ty, int32_t *threads) {
    if (!s) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null solver");
    if (tx) *tx = s->ts.tx;
    if (ty) *ty = s->ts.ty;
    if (threads) *threads = s->ts.nt;
    return CDC_OK;
}

int32_t cdc_solver_get_tile(cdc_solver_t *s, int32_t *tx, int32_t *ty) {
    if (!s) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null solver");
    if (tx) *tx = s->ts.tx;
    if (ty) *ty = s->ts.ty;
    return CDC_OK;
}

int32_t cdc_solver_autotune(cdc_solver_t *s, int32_t reps, int32_t *tx, int32_t *ty, double *best_ms) {
    if (!s) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null solver");
    if (reps < 1) reps = 1;
    const cdc::TileShape *shapes;
    const int n = cdc::tile_shapes(&shapes);
    const double L = 2.0 * s->mesh.domain_radius / (double)s->nb;
    double el[4];
    cdc::orbital_elements(s->phys, el);
    const int scratch = (s->cur + 1) % NBUF;
    double best = 1e300;
    cdc::TileShape pick = s->ts;
    std::vector<hipEvent_t> marks(2 * ((size_t)reps + 1), nullptr);
    for (auto &m : marks) HIP_TRY(hipEventCreate(&m));
    int rc = CDC_OK;
    // the marching shapes read the committed state's strip-edge columns
    if (s->d_edge[0] && s->edge_version != s->version) {
        HIP_TRY(cdc::launch_edge_build(s->nv, s->N, s->d_u[s->cur], s->d_edge[s->cur], s->stream));
        s->edge_version = s->version;
    }
    for (int i = 0; i < n && rc == CDC_OK; ++i) {
        const cdc::TileShape t = shapes[i];
        if (s->bs % t.tx != 0 || s->bs % t.ty != 0) continue;
        if (cdc::march_shape(t) && !cdc::march_fits(s->nv, s->N)) continue;
        // stage launches only (no finalize) from the committed state into a
        // scratch buffer, in the costlier later-stage form
        cdc::StageArgs a;
        memset(&a, 0, sizeof(a));
        a.uc = s->d_u[s->cur];
        a.u0 = s->d_init;
        a.stat = s->d_stat;
        a.ub = s->d_u[s->cur];
        a.aux = s->d_aux;
        a.uo = s->d_u[scratch];
        a.ec = s->d_edge[s->cur];   // null when no shape marches; eo stays null (not written)
        // the tile-sum buffer that does not hold the committed max speed
        a.tile_sums = s->ts_last == s->d_tile_sums ? s->d_tile_sums + s->ts_stride : s->d_tile_sums;
        a.tile_err = s->d_tile_err;
        a.err_flag = s->d_err_flag;
        a.N = s->N;
        a.nb = s->nb;
        a.bs = s->bs;
        a.tpb = (s->bs / t.tx) * (s->bs / t.ty);
        a.ntiles = s->nblk * a.tpb;
        a.stage_id = 1;
        a.read_base = 1;
        a.reuse_prim = s->hydro.has_density_floor ? 0 : 1;
        a.R = s->mesh.domain_radius;
        a.L = L;
        a.time = s->time;
        a.dt = 1e-9;
        a.c1 = 0.5;
        a.c0 = 0.5;
        a.orb = cdc::orbital_state(el, s->time);
        a.k = make_consts(s, s->time, a.dt);
        const std::vector<cdc::StageArgs> run((size_t)reps + 1, a);   // first launch warms up
        if ((rc = cdc::host::launch_set_int_checked(s))) break;
        if ((rc = cdc::host::launch_stages(s, run.data(), run.size(), t, marks.data(), false))) break;
        const hipError_t es = hipEventSynchronize(marks.back());
        if (es != hipSuccess) { rc = set_msg(CDC_ERR_HIP, hipGetErrorString(es)); break; }
        double sum = 0.0;
        for (int r = 1; r <= reps; ++r) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, marks[2 * r], marks[2 * r + 1]);
            sum += ms;
        }
        const double per = sum / reps;
        if (per < best) { best = per; pick = t; }
    }
    for (auto &m : marks) (void)hipEventDestroy(m);
    if (rc) return rc;
    HIP_TRY(cdc::launch_set_int(s->d_err_flag, 0, s->stream));
    rc = cdc_solver_set_kernel(s, pick.tx, pick.ty, pick.nt);
    if (rc) return rc;
    if (tx) *tx = pick.tx;
    if (ty) *ty = pick.ty;
    if (best_ms) *best_ms = best;
    return CDC_OK;
}

int32_t cdc_solver_max_signal_speed(cdc_solver_t *s, double *out) {
    if (!s || !out) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null pointer");
    return compute_amax(s, out);
}

int32_t cdc_solver_set_timing(cdc_solver_t *s, int32_t enable) {
    if (!s) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null solver");
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->timing = enable != 0;
    s->launches = 0;
    s->total_ms = 0.0;
    return CDC_OK;
}

// A copy of a solver's committed state in device memory (the fallback
// states of the driver loop, src/main.rs:148-200, without a host round trip).
struct cdc_snapshot {
    int nv = 0, nblk = 0;
    size_t ncell = 0;
    double *u = nullptr, *ist = nullptr, *oec = nullptr, *tq = nullptr;
    double time = 0.0;
    int64_t num = 0, den = 1;
    bool valid = false;
};

int32_t cdc_snapshot_create(const cdc_solver_t *s, cdc_snapshot_t **out) {
    if (!s || !out) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null pointer");
    *out = nullptr;
    cdc_snapshot *p = new cdc_snapshot();
    p->nv = s->nv;
    p->nblk = s->nblk;
    p->ncell = s->ncell();
    if (hipMalloc((void **)&p->u, s->nstate() * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&p->ist, (size_t)s->nblk * 7 * s->nv * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&p->oec, (size_t)s->nblk * 7 * 4 * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&p->tq, (size_t)s->nblk * 2 * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        cdc_snapshot_destroy(p);
        return set_msg(CDC_ERR_OUT_OF_MEMORY, "snapshot buffers");
    }
    *out = p;
    return CDC_OK;
}

void cdc_snapshot_destroy(cdc_snapshot_t *p) {
    if (!p) return;
    (void)hipFree(p->u);
    (void)hipFree(p->ist);
    (void)hipFree(p->oec);
    (void)hipFree(p->tq);
    delete p;
}

static bool snapshot_fits(const cdc_solver *s, const cdc_snapshot *p) {
    return p->nv == s->nv && p->nblk == s->nblk && p->ncell == s->ncell();
}

int32_t cdc_solver_save(cdc_solver_t *s, cdc_snapshot_t *p) {
    if (!s || !p) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null pointer");
    if (!snapshot_fits(s, p)) return set_msg(CDC_ERR_INVALID_ARGUMENT, "snapshot made for another mesh");
    const int c = s->cur;
    HIP_TRY(hipMemcpyAsync(p->u, s->d_u[c], s->nstate() * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(p->ist, s->d_ist[c], (size_t)s->nblk * 7 * s->nv * sizeof(double),
                           hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(p->oec, s->d_oec[c], (size_t)s->nblk * 7 * 4 * sizeof(double), hipMemcpyDeviceToDevice,
                           s->stream));
    HIP_TRY(hipMemcpyAsync(p->tq, s->d_tq[c], (size_t)s->nblk * 2 * sizeof(double), hipMemcpyDeviceToDevice,
                           s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    p->time = s->time;
    p->num = s->num;
    p->den = s->den;
    p->valid = true;
    return CDC_OK;
}

int32_t cdc_solver_restore(cdc_solver_t *s, const cdc_snapshot_t *p) {
    if (!s || !p) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null pointer");
    if (!snapshot_fits(s, p)) return set_msg(CDC_ERR_INVALID_ARGUMENT, "snapshot made for another mesh");
    if (!p->valid) return set_msg(CDC_ERR_INVALID_ARGUMENT, "snapshot holds no state");
    const int c = (s->cur + 1) % NBUF;
    HIP_TRY(hipMemcpyAsync(s->d_u[c], p->u, s->nstate() * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_ist[c], p->ist, (size_t)s->nblk * 7 * s->nv * sizeof(double),
                           hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_oec[c], p->oec, (size_t)s->nblk * 7 * 4 * sizeof(double), hipMemcpyDeviceToDevice,
                           s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_tq[c], p->tq, (size_t)s->nblk * 2 * sizeof(double), hipMemcpyDeviceToDevice,
                           s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->cur = c;
    s->time = p->time;
    s->num = p->num;
    s->den = p->den;
    s->amax_valid = false;
    s->version += 1;
    return CDC_OK;
}

int32_t cdc_snapshot_time(const cdc_snapshot_t *p, double *time) {
    if (!p || !time) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null pointer");
    *time = p->time;
    return CDC_OK;
}

int32_t cdc_solver_get_torque(cdc_solver_t *s, double *torque) {
    if (!s || !torque) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null pointer");
    HIP_TRY(hipMemcpyAsync(torque, s->d_tq[s->cur], (size_t)s->nblk * 2 * sizeof(double), hipMemcpyDeviceToHost,
                           s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return CDC_OK;
}

int32_t cdc_solver_set_torque(cdc_solver_t *s, const double *torque) {
    if (!s || !torque) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null pointer");
    HIP_TRY(hipMemcpyAsync(s->d_tq[s->cur], torque, (size_t)s->nblk * 2 * sizeof(double), hipMemcpyHostToDevice,
                           s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->version += 1;
    return CDC_OK;
}

int32_t cdc_solver_time_series_sample(cdc_solver_t *s, double *time, double *ist_total, double *oec_total,
                                      double *torque_total) {
    if (!s) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null solver");
    const size_t ni = (size_t)s->nblk * 7 * s->nv, ne = (size_t)s->nblk * 7 * 4, nt = (size_t)s->nblk * 2;
    std::vector<double> ist(ni), oec(ne), tq(nt);
    HIP_TRY(hipMemcpyAsync(ist.data(), s->d_ist[s->cur], ni * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(oec.data(), s->d_oec[s->cur], ne * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(tq.data(), s->d_tq[s->cur], nt * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (time) *time = s->time;
    // sums over blocks in Mesh::block_indexes() order
    if (ist_total)
        for (size_t k = 0; k < (size_t)7 * s->nv; ++k) {
            double a = 0.0;
            for (int b = 0; b < s->nblk; ++b) a += ist[(size_t)b * 7 * s->nv + k];
            ist_total[k] = a;
        }
    if (oec_total)
        for (size_t k = 0; k < 7 * 4; ++k) {
            double a = 0.0;
            for (int b = 0; b < s->nblk; ++b) a += oec[(size_t)b * 7 * 4 + k];
            oec_total[k] = a;
        }
    if (torque_total)
        for (size_t k = 0; k < 2; ++k) {
            double a = 0.0;
            for (int b = 0; b < s->nblk; ++b) a += tq[(size_t)b * 2 + k];
            torque_total[k] = a;
        }
    return CDC_OK;
}

int32_t cdc_solver_get_timing(cdc_solver_t *s, int64_t *launches, double *total_ms, double *cells_per_launch) {
    if (!s) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null solver");
    if (launches) *launches = s->launches;
    if (total_ms) *total_ms = s->total_ms;
    if (cells_per_launch) *cells_per_launch = (double)s->ncell();
    return CDC_OK;
}

static int advance_at(cdc_solver_t *s, int32_t fold, double dt, cdc_error_t *err) {
    int rc = CDC_OK;
    StepPlan plan;
    cdc::host::plan_steps(s, fold, dt, plan);
    const size_t nst = plan.args.size();
    // the last stage's finalize leaves the result's max signal speed (the
    // next call's dt) in d_amax_adv; with beta cooling the stages read the
    // advance's relaxation factor from aux plane 2
    plan.args.back().amax_out = s->d_amax_adv;
    for (auto &a : plan.args) a.cool_factor = s->naux > 2 ? 1 : 0;
    if (s->timing) {
        while (s->ev.size() < 2 * nst) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            s->ev.push_back(e);
        }
    }
    rc = cdc::host::stage_args(s, plan.args.data(), nst);
    if (rc) return rc;
    if ((rc = cdc::host::ensure_edges(s, s->stream))) return rc;
    rc = cdc::host::enqueue_advance(s, nst, plan.args[0].ntiles, s->timing ? s->ev.data() : nullptr);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    int32_t flag = *s->h_flag;
    if (flag != 0 && nst > 1) {
        // The fused launches may have run stages after a failed finalize;
        // repeat the call unfused from the untouched committed state, where
        // a failed stage skips the rest, so that the flag and the records
        // name the first error as the reference reports it.
        HIP_TRY(cdc::launch_reset_advance(s->d_err_flag, s->d_amax_adv, s->stream));
        rc = cdc::host::launch_stages(s, plan.args.data(), nst, s->ts, nullptr, true, false);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(s->h_flag, s->d_err_flag, sizeof(int32_t), hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipMemcpyAsync(s->h_amax, s->d_amax_adv, sizeof(double), hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        flag = *s->h_flag;
    }
    if (s->timing) {
        for (size_t i = 0; i < nst; ++i) {
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, s->ev[2 * i], s->ev[2 * i + 1]));
            s->total_ms += ms;
            s->launches += 1;
        }
    }
    if (flag == 0) {
        cdc::host::commit(s, plan);
        s->amax_host = *s->h_amax;
        s->amax_host_version = s->version;
        return CDC_OK;
    }
    return cdc::host::resolve_error(s, flag, plan, err);
}

int32_t cdc_solver_advance(cdc_solver_t *s, int32_t fold, double *dt_out, cdc_error_t *err) {
    if (err) memset(err, 0, sizeof(*err));
    if (!s) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null solver");
    if (fold < 0) return set_msg(CDC_ERR_INVALID_ARGUMENT, "fold must be >= 0");

    // dt = cfl * dx / max signal speed (src/scheme.rs:336-340)
    double amax = 0.0;
    int rc = compute_amax(s, &amax);
    if (rc) return rc;
    const double dt = cdc::host::step_dt(s, amax);
    if (dt_out) *dt_out = dt;
    if (fold == 0) return CDC_OK;
    return advance_at(s, fold, dt, err);
}

int32_t cdc_solver_advance_dt(cdc_solver_t *s, int32_t fold, double dt, cdc_error_t *err) {
    if (err) memset(err, 0, sizeof(*err));
    if (!s) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null solver");
    if (fold < 0) return set_msg(CDC_ERR_INVALID_ARGUMENT, "fold must be >= 0");
    if (!(dt > 0.0) || !std::isfinite(dt)) return set_msg(CDC_ERR_INVALID_ARGUMENT, "dt must be positive and finite");
    if (fold == 0) return CDC_OK;
    return advance_at(s, fold, dt, err);
}

int32_t cdc_solver_resample_from(cdc_solver_t *dst, const cdc_solver_t *src) {
    if (!dst || !src) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null solver");
    if (dst == src) return set_msg(CDC_ERR_INVALID_ARGUMENT, "source and destination are one solver");
    if (dst->nv != src->nv) return set_msg(CDC_ERR_INVALID_ARGUMENT, "solvers must share one hydro kind");
    if (dst->mesh.domain_radius != src->mesh.domain_radius)
        return set_msg(CDC_ERR_INVALID_ARGUMENT, "meshes must cover one domain");
    if (dst->device != src->device) return set_msg(CDC_ERR_INVALID_ARGUMENT, "solvers must share one device");
    const int64_t Ns = src->N, Nd = dst->N;
    const bool down = Ns >= Nd;
    const int64_t r = down ? Ns / Nd : Nd / Ns;
    if ((down ? Ns % Nd : Nd % Ns) != 0 || (r & (r - 1)) != 0 || r > 16)
        return set_msg(CDC_ERR_INVALID_ARGUMENT, "grid sizes must differ by a power-of-two factor of at most 16");
    HIP_TRY(hipStreamSynchronize(src->stream));
    // write into a free buffer of dst and commit it: a failure leaves dst as it was
    const int out = (dst->cur + 1) % NBUF;
    HIP_TRY(cdc::launch_resample(dst->nv, src->d_u[src->cur], Ns, dst->d_u[out], Nd, (int)r, down ? 1 : 0,
                                 dst->stream));
    const size_t ni = (size_t)src->nblk * 7 * src->nv, ne = (size_t)src->nblk * 7 * 4;
    const size_t di = (size_t)dst->nblk * 7 * dst->nv, de = (size_t)dst->nblk * 7 * 4;
    std::vector<double> is(ni), es(ne), id(di), ed(de);
    HIP_TRY(hipMemcpy(is.data(), src->d_ist[src->cur], ni * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(es.data(), src->d_oec[src->cur], ne * sizeof(double), hipMemcpyDeviceToHost));
    cdc::host::resample_block_sums(&src->mesh, &dst->mesh, dst->nv, is.data(), es.data(), id.data(), ed.data());
    HIP_TRY(hipMemcpyAsync(dst->d_ist[out], id.data(), di * sizeof(double), hipMemcpyHostToDevice, dst->stream));
    HIP_TRY(hipMemcpyAsync(dst->d_oec[out], ed.data(), de * sizeof(double), hipMemcpyHostToDevice, dst->stream));
    HIP_TRY(hipMemsetAsync(dst->d_tq[out], 0, (size_t)dst->nblk * 2 * sizeof(double), dst->stream));
    HIP_TRY(hipStreamSynchronize(dst->stream));
    dst->cur = out;
    dst->time = src->time;
    dst->num = src->num;
    dst->den = src->den;
    dst->amax_valid = false;
    dst->version += 1;
    return CDC_OK;
}

}  // extern "C"
