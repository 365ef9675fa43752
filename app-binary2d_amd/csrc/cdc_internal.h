// cdc_internal.h -- launch interface between the host driver (cdc_solver.cpp)
// and the gfx950 kernels (cdc_kernels.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "../../include/cdc_solver.h"
#include "cdc_orbit.h"

namespace cdc {

// Per-tile error record of the fused stage kernel: the first failing cell of
// the tile in block-local C order (cell = i * bs + j).
struct TileErr {
    int32_t code;
    int32_t cell;
    double value;
};

// Physics constants of a stage launch, folded once on the host.
struct StageConsts {
    double dx, dy, idx, idy;
    double rs2;                 // softening_length^2
    double inv_mach2;           // isothermal: 1 / Mach^2
    double gm, gm1, inv_gm1;    // Euler gamma, gamma - 1, 1 / (gamma - 1)
    double plm, pdt;            // PLM theta, plm_density_trigger (0 when unset)
    double visc;                // nu (constant_nu) or alpha / sqrt(gamma) (alpha)
    double lam;                 // lambda_multiplier
    double eps_dx;              // 1e-12 * dx (alpha model)
    double sink_s2, inv_sink_s2, sink_rate;
    double fake_threshold, fake_rate;
    double cooling_coefficient, cooling_inv_beta;
    double cooling_f;           // T^4 slow-start factor exp(-2 pi / (t + dt)) or 1
    double mach_ceiling_coef;   // 2 / gm / (gm - 1) / mach_ceiling^2
    double density_floor, pressure_floor, density_floor_pressure;
    // the floors as compared: the floor when it is set, -inf otherwise (no
    // state compares below -inf), so that no flag is tested per cell
    double density_floor_cmp, pressure_floor_cmp;
    int32_t viscosity_model, sink_model, one_body, has_mach_ceiling;
    int32_t has_density_floor, has_pressure_floor, density_index, cooling_slow_start;
    int32_t cooling_model, has_pdt;
};

// Static per-cell planes, SoA [plane][gi][gj], written once per physics:
// plane 0 = buffer-zone rate 0.5 rate (1 + tanh((r - R)/scale)) Omega_outer
// (Physics::source_terms, src/physics.rs:443-446); plane 1 (beta cooling
// only) = Omega = sqrt(M / (r^2 + r_s^2)^(3/2)).
struct AuxArgs {
    double *aux;
    int32_t N, nb, bs, naux;
    double R, L, buffer_rate, buffer_scale, omega_outer, rs2, mtot;
};

struct StageArgs {
    const double *uc;      // stage input (device state layout, sidx)
    const double *u0;      // initial (background) state, the same layout
    const double *ub;      // RK step base (read when read_base)
    const double *aux;     // static per-cell planes
    double *uo;            // stage output
    double *tile_sums;     // [ntiles][NSUM + 1] (last slot: max signal speed)
    TileErr *tile_err;     // [ntiles]
    int32_t *err_flag;     // stage id of the first failing stage, 0 if none
    // per-block integrated_source_terms / orbital_elements_change
    const double *ist_c, *ist_b;   // [nblk][7][nv]
    const double *oec_c, *oec_b;   // [nblk][7][4]
    double *ist_o, *oec_o;
    // per-block torque impulses on body 1, 2 about the origin ([nblk][2])
    const double *tq_c, *tq_b;
    double *tq_o;
    int32_t *perr;                 // [nblk] perturbation error of the block
    // when set (the last stage of an advance), the finalize also folds the
    // blocks' max signal speeds into *amax_out (zeroed beforehand)
    double *amax_out;
    // the strip-edge columns of uc and uo (column-marching kernel; null when
    // the block size is not a multiple of 64): per plane, row and boundary
    // k between 64-column strips (global column 64 k), the four columns
    // 64 k - 2 .. 64 k + 1 (periodic), [nv][N][N / 64][4]; the kernel reads
    // ec and writes eo
    const double *ec;
    double *eo;
    // the column-marching kernel's static operands in 16-byte pairs, one
    // pair plane [N][N][2] after another: (u0[0], u0[1]), (u0[2], u0[3] or,
    // for 3 variables, the buffer rate) and, with beta cooling, (buffer
    // rate, the advance's cooling factor); null when no strip shape fits
    const double *stat;
    int32_t N, nb, bs, ntiles, tpb;  // tpb: tiles per block
    int32_t stage_id, read_base, want_amax, reuse_prim;
    int32_t cool_factor;   // beta cooling reads exp(-dt Omega / beta) from aux plane 2
    double R, L;           // domain radius, block length
    double time, dt, c1, c0, area;
    double el[4];          // Physics::orbital_elements
    Orbit orb;
    StageConsts k;
};

// The cooling law every run of a stage launch shares, for kernels compiled
// for one: T^4, beta cooling reading exp(-dt Omega / beta) from aux plane 2
// (cool_factor), beta cooling evaluating it; kCoolAny: isothermal gas (no
// cooling), and the tiled kernel, which reads the law from each run.
enum StageCooling { kCoolAny = 0, kCoolT4 = 1, kCoolBetaFactor = 2, kCoolBetaExp = 3 };
// The viscosity law every run of a stage launch shares, for the column-
// marching kernel (constant nu includes nu = 0).
enum StageViscosity { kViscConst = 1, kViscAlpha = 2 };

// Tile shape chosen for a block size (the largest shape that divides it).
struct TileShape {
    int tx, ty, nt;   // tile of tx x ty cells, nt threads
};
TileShape tile_shape_for(int bs, int nv);
// the shapes of the column-marching kernel: 64-column strips, one wave each
inline bool march_shape(const TileShape &ts) { return ts.ty == 64 && ts.nt == 64; }
// whether the column-marching kernel can run an N x N grid of nv conserved
// variables: it addresses its arrays through 32-bit buffer offsets, so each
// (nv state planes, up to 3 aux planes) must hold under 2^32 bytes
inline bool march_fits(int nv, int64_t N) {
    return (uint64_t)(nv > 3 ? nv : 3) * (uint64_t)N * (uint64_t)N * 8u < (1ull << 32);
}
int tile_shapes(const TileShape **out);

int nsum_for(int nv);

hipError_t launch_aux(const AuxArgs &a, hipStream_t s);

// One RK stage (fused stage kernel, then the per-block finalize) over a batch
// of runs: device arrays of StageArgs, one per run; tile_begin / block_begin
// hold nruns + 1 prefix offsets.  A single run is a batch of one.  With fin
// (the earlier stage's per-run arguments), the launch also finalizes that
// stage over the nblk blocks that block_begin divides among the runs, in
// workgroups ahead of the tiles.  ev0 / ev1 (either may be null) take the
// launch's start and end.  `cool` (StageCooling), `visc` (StageViscosity)
// and `read_base` (whether the stages read the step base, the same for every
// run) let the launch pick a kernel compiled for them; the tiled kernel
// decides them at run time.
hipError_t launch_stage_batch(int nv, const TileShape &ts, const StageArgs *args, const int32_t *tile_begin,
                              int nruns, int ntiles, const StageArgs *fin, const int32_t *block_begin, int nblk,
                              int cool, int visc, int read_base, hipEvent_t ev0, hipEvent_t ev1, hipStream_t s);
// The column-marching stage kernel (cdc_march.hip): strips of 64 columns x
// `rows` rows (32 or 16), one wave each; otherwise as launch_stage_batch.
hipError_t launch_march_batch(int nv, int rows, int cool, int visc, int read_base, const StageArgs *args,
                              const int32_t *begin, int nruns, int ntiles, const StageArgs *fin,
                              const int32_t *fin_begin, int fin_nruns, int nblk, hipEvent_t ev0, hipEvent_t ev1,
                              hipStream_t s);
// The strip-edge columns of the state u (the device state layout, sidx) into E (the layout
// of StageArgs::ec), N a multiple of 64.
hipError_t launch_edge_build(int nv, int N, const double *u, double *E, hipStream_t s);
// StageArgs::stat from the initial state u0 (the device state layout) and the aux
// planes (buffer rate in plane 0); npair = 3 leaves the cooling-factor slot
// of the third pair plane to launch_cooling_factor.
hipError_t launch_stat_build(int nv, int npair, const double *u0, const double *aux, double *stat, int64_t nn,
                             hipStream_t s);
hipError_t launch_finalize_batch(int nv, const StageArgs *args, const int32_t *block_begin, int nruns,
                                 int nblocks, hipStream_t s);
hipError_t launch_batch_reset(const StageArgs *args, int nruns, hipStream_t s);
hipError_t launch_batch_tail(int nv, const StageArgs *args, int nruns, double *amax, int32_t *flags,
                             hipStream_t s);
hipError_t launch_max_speed(int nv, const double *u, int64_t ncell, double gm, double *partial,
                            int nparts, hipStream_t s);
hipError_t launch_aos_to_soa(int nv, int nb, int bs, const double *aos, double *soa, hipStream_t s);
hipError_t launch_soa_to_aos(int nv, int nb, int bs, const double *soa, double *aos, hipStream_t s);
hipError_t launch_set_int(int32_t *p, int32_t v, hipStream_t s);
// Zeroes an advance's error flag and max-speed slot (a kernel, so that a
// captured graph holds no memset node).
hipError_t launch_reset_advance(int32_t *flag, double *amax, hipStream_t s);
// exp(-dt Omega / beta) from aux plane 1 and a->dt (nn cells): into the
// cooling slot of stat's third pair plane when stat is not null (the stages
// march), else into aux plane 2
hipError_t launch_cooling_factor(const StageArgs *a, double *aux, double *stat, int64_t nn, hipStream_t s);
// out[3 i .. 3 i + 2] = the stage kernel's 1/x, 1/sqrt(x), sqrt(x) of x[i]
hipError_t launch_math_probe(const double *x, double *out, int n, hipStream_t s);
// r in 1..16 (a power of two), down = 1 for averaging, 0 for injection
hipError_t launch_resample(int nv, const double *src, int64_t Ns, double *dst, int64_t Nd, int r, int down,
                           hipStream_t s);
hipError_t launch_column_max(const double *x, int n, int stride, double *out, hipStream_t s);

}  // namespace cdc
