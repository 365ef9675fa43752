/*
 * cdc_solver.h -- C ABI of the MI355X-native per-block Runge-Kutta update of
 * the circumbinary disk code (CDC, clemson-cal/app-binary2d).
 *
 * The entry points below replace the reference's hot path
 *
 *     pub fn advance<H>(state, hydro, mesh, physics, fold, &mut timestep,
 *                       block_data, runtime) -> Result<State, HydroError>
 *                                                       (src/scheme.rs:323-349)
 *
 * and the data it moves (State / BlockState, src/state.rs:27-62).  A Rust
 * caller binds them with an `extern "C"` block over these declarations.
 *
 * Data layout at the boundary (identical to the reference's in-memory arrays,
 * so a checkpoint converts with a memcpy):
 *
 *   conserved : [block b][i][j][v]   doubles, b in Mesh::block_indexes() order
 *               (i-major, src/mesh.rs:94-99), i = x index, j = y index
 *               (ndarray C order, src/state.rs:46), v < nv
 *               (nv = 3 isothermal, 4 Euler)
 *   ist       : [block b][item k][v]  integrated_source_terms, the 7 items in
 *               ItemizedChange field order (src/state.rs:28-36):
 *               sink1, sink2, grav1, grav2, buffer, cooling, fake_mass
 *   oec       : [block b][item k][4]  orbital_elements_change, each an
 *               OrbitalElements tuple (a, M, q, e) (src/physics.rs:411-417)
 *
 * Every call is synchronous with respect to the host.  A handle owns one HIP
 * stream and its device buffers; a handle is not re-entrant, distinct handles
 * may be used from distinct threads.  Host buffers are caller-owned and are
 * copied.
 */
#ifndef CDC_SOLVER_H
#define CDC_SOLVER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CDC_ABI_VERSION 1
#define CDC_NUM_ITEMS 7

/* ItemizedChange field order, src/state.rs:28-36 */
enum {
    CDC_ITEM_SINK1 = 0,
    CDC_ITEM_SINK2 = 1,
    CDC_ITEM_GRAV1 = 2,
    CDC_ITEM_GRAV2 = 3,
    CDC_ITEM_BUFFER = 4,
    CDC_ITEM_COOLING = 5,
    CDC_ITEM_FAKE_MASS = 6
};

/* Return codes.  1..3 mirror HydroErrorType (src/physics.rs:26-36). */
enum {
    CDC_OK = 0,
    CDC_ERR_NEGATIVE_DENSITY = 1,
    CDC_ERR_NEGATIVE_PRESSURE = 2,
    CDC_ERR_UNBOUND_ORBIT = 3,
    CDC_ERR_INVALID_ARGUMENT = 100,
    CDC_ERR_HIP = 101,
    CDC_ERR_OUT_OF_MEMORY = 102,
    CDC_ERR_NO_DEVICE = 103
};

enum { CDC_ISOTHERMAL = 0, CDC_EULER = 1 };                 /* AnyHydro, src/app.rs:28-33 */
enum { CDC_VISC_CONSTANT_NU = 0, CDC_VISC_ALPHA = 1 };      /* ViscosityModel, src/physics.rs:100-103 */
enum { CDC_SINK_IMPARTIAL = 0, CDC_SINK_TORQUE_FREE = 1 };  /* SinkModel, src/physics.rs:112-115 */
/* Cooling law of the Euler system.  T4 is the reference's
 * (src/physics.rs:714-739); BETA (t_cool = beta / Omega, relaxing the
 * specific internal energy to that of the initial state) has no reference
 * counterpart. */
enum { CDC_COOLING_T4 = 0, CDC_COOLING_BETA = 1 };

/* Mesh, src/mesh.rs:19-37 */
typedef struct cdc_mesh {
    int32_t num_blocks;
    int32_t block_size;
    double  domain_radius;
} cdc_mesh_t;

/* AnyHydro: Isothermal (src/physics.rs:258-265) | Euler (src/physics.rs:273-303).
 * Option<f64> fields are a has_/value pair. */
typedef struct cdc_hydro {
    int32_t kind;                 /* CDC_ISOTHERMAL | CDC_EULER */
    int32_t density_index;        /* Euler; reference default 2 */
    double  mach_number;          /* Isothermal */
    double  gamma_law_index;      /* Euler */
    double  cooling_coefficient;  /* Euler, T^4 law */
    int32_t cooling_slow_start;   /* Euler, bool */
    int32_t has_density_floor;
    double  density_floor;
    int32_t has_pressure_floor;   /* Euler */
    int32_t cooling_model;        /* CDC_COOLING_T4 | CDC_COOLING_BETA */
    double  pressure_floor;
    double  cooling_beta;         /* CDC_COOLING_BETA only */
} cdc_hydro_t;

/* Physics, src/physics.rs:123-217.  `viscosity_model`/`viscosity` hold the
 * resolved Physics::viscosity_model() (src/physics.rs:222-230): the caller
 * applies the precedence of the deprecated `nu` / `alpha` fields. */
typedef struct cdc_physics {
    double  buffer_rate;
    double  buffer_scale;
    double  binary_eccentricity;
    double  binary_mass_ratio;
    double  lambda_multiplier;
    int32_t viscosity_model;      /* CDC_VISC_* */
    int32_t sink_model;           /* CDC_SINK_* */
    double  viscosity;            /* nu (constant_nu) or alpha */
    double  cfl;
    double  plm;
    int32_t has_plm_density_trigger;
    int32_t one_body;             /* bool */
    double  plm_density_trigger;
    double  fake_mass_threshold;
    double  fake_mass_rate;
    int32_t has_mach_ceiling;
    int32_t rk_order;             /* 1, 2 or 3 */
    double  mach_ceiling;
    double  sink_radius;
    double  sink_rate;
    int32_t has_softening_length;
    int32_t has_dt_min;
    double  softening_length;
    double  dt_min;
} cdc_physics_t;

/* kepler_two_body::PointMass field order (mass, x, y, vx, vy) */
typedef struct cdc_point_mass {
    double mass;
    double position_x;
    double position_y;
    double velocity_x;
    double velocity_y;
} cdc_point_mass_t;

/* HydroError, src/physics.rs:48-71: error type + value, position, and the
 * binary's orbital state at the failing stage's time. */
typedef struct cdc_error {
    int32_t code;                 /* CDC_ERR_* (0 when no error) */
    int32_t has_binary;
    double  value;                /* NegativeDensity / NegativePressure payload */
    double  x;
    double  y;
    double  time;                 /* start time of the failing RK stage */
    cdc_point_mass_t binary[2];
    char    message[256];
} cdc_error_t;

typedef struct cdc_solver cdc_solver_t;

/* ---- library ------------------------------------------------------------ */

/* ABI version (CDC_ABI_VERSION). */
int32_t cdc_abi_version(void);

/* sizeof of each boundary struct, so that bindings can verify their layout. */
int32_t cdc_struct_size(int32_t which);  /* 0 mesh, 1 hydro, 2 physics, 3 error,
                                            4 cavity model, 5 closure config */

/* Number of visible HIP devices (0 when none). */
int32_t cdc_device_count(void);

/* The stage kernel's FP64 1/x, 1/sqrt(x) and sqrt(x) (hardware estimates
 * refined by one Newton step) of n host values on the current device:
 * out[3 i + 0 .. 2] for x[i].  No reference counterpart: it exposes the
 * device arithmetic so that its error bounds can be measured. */
int32_t cdc_math_probe(const double *x, double *out, int32_t n);

/* Static text for a return code. */
const char *cdc_error_string(int32_t code);

/* Message of the last failing call on this thread. */
const char *cdc_last_error_message(void);

/* ---- solver handle: BlockData + State + advance ------------------------- */

/*
 * Create a solver.  `initial_conserved` ([block][i][j][v], see above) plays the
 * role of BlockData::initial_conserved (src/scheme.rs:36-58, used by the buffer
 * and fake-mass sources); it also becomes the initial state at time 0,
 * iteration 0/1 with zero ist/oec (State::from_model, src/state.rs:99-110).
 */
int32_t cdc_solver_create(const cdc_mesh_t *mesh, const cdc_hydro_t *hydro,
                          const cdc_physics_t *physics,
                          const double *initial_conserved, cdc_solver_t **out);

void cdc_solver_destroy(cdc_solver_t *solver);

/* Replace the physics used by later advance calls (the safe-mode swap of
 * src/main.rs:131-147 passes a modified Physics per call). */
int32_t cdc_solver_set_physics(cdc_solver_t *solver, const cdc_physics_t *physics);

/* Upload a State (src/state.rs:57-62).  ist / oec may be NULL (zeros). */
int32_t cdc_solver_set_state(cdc_solver_t *solver, double time,
                             int64_t iteration_num, int64_t iteration_den,
                             const double *conserved, const double *ist,
                             const double *oec);

/* Download the current State.  Any output pointer may be NULL. */
int32_t cdc_solver_get_state(cdc_solver_t *solver, double *time,
                             int64_t *iteration_num, int64_t *iteration_den,
                             double *conserved, double *ist, double *oec);

/*
 * scheme::advance (src/scheme.rs:323-349): dt = cfl * dx / max signal speed
 * (raised to dt_min when set), written to *dt_out, then `fold` RK steps of
 * that dt.  On a hydro error the call returns its code, fills *err (may be
 * NULL) and leaves the handle's state exactly as it was before the call.
 */
int32_t cdc_solver_advance(cdc_solver_t *solver, int32_t fold, double *dt_out,
                           cdc_error_t *err);

/*
 * The same `fold` RK steps at a caller-given dt (no CFL evaluation; the
 * reference always derives dt, src/scheme.rs:336-340).  For building pairs of
 * runs at one time step, e.g. a coarse step of dt against 4 fine steps of
 * dt/4.  Errors and transactionality as cdc_solver_advance.
 */
int32_t cdc_solver_advance_dt(cdc_solver_t *solver, int32_t fold, double dt,
                              cdc_error_t *err);

/*
 * Re-grid the committed state of `src` into `dst` (one domain, one hydro
 * kind, grid sizes differing by a power-of-two factor of at most 16), on the
 * device: injection when dst is finer (the checkpoint up-sampler,
 * upsampler/src/main.rs:27-72), the mean of r x r cells when it is coarser
 * (pairwise-tree sums, so down(up(u)) == u exactly).  dst takes src's time
 * and iteration; its ist / oec follow cdc_resample_state.  dst's initial
 * state (the buffer-zone target) is not changed.
 */
int32_t cdc_solver_resample_from(cdc_solver_t *dst, const cdc_solver_t *src);

/*
 * Per-block torque impulses on body 1 and 2 about the origin, [block][2]:
 * the sum over stages of x_k x (-dp_k), dp_k the block's gas momentum change
 * from body k's gravity (the grav items) and x_k body k's position at the
 * stage's time, carried through the RK combine like integrated_source_terms.
 * No counterpart in the reference (its torque is derived from time series
 * offline); zeroed by cdc_solver_set_state.
 */
int32_t cdc_solver_get_torque(cdc_solver_t *solver, double *torque);

/* Restore per-block torque impulses ([block][2], as cdc_solver_get_torque
 * returns them) into the current state, after cdc_solver_set_state, so that
 * a run restarted from a checkpoint carries its running totals on.  No
 * reference counterpart (the reference keeps no torque accumulator). */
int32_t cdc_solver_set_torque(cdc_solver_t *solver, const double *torque);

/*
 * State::time_series_sample (src/state.rs:135-146): the time and the sums
 * over blocks (in block order) of integrated_source_terms [7][nv] and
 * orbital_elements_change [7][4], plus the summed torque impulses [2].  Any
 * output may be NULL.
 */
int32_t cdc_solver_time_series_sample(cdc_solver_t *solver, double *time, double *ist_total,
                                      double *oec_total, double *torque_total);

/*
 * Device-resident copies of the committed state (conserved, ist, oec,
 * torque, time, iteration): the fallback states the driver loop keeps
 * (src/main.rs:148-200 clones the State per advance), copied at device
 * bandwidth instead of through the host.  A snapshot fits solvers of the
 * mesh and hydro kind it was created for.
 */
typedef struct cdc_snapshot cdc_snapshot_t;
int32_t cdc_snapshot_create(const cdc_solver_t *solver, cdc_snapshot_t **out);
void    cdc_snapshot_destroy(cdc_snapshot_t *snapshot);
int32_t cdc_snapshot_time(const cdc_snapshot_t *snapshot, double *time);
int32_t cdc_solver_save(cdc_solver_t *solver, cdc_snapshot_t *snapshot);
int32_t cdc_solver_restore(cdc_solver_t *solver, const cdc_snapshot_t *snapshot);

/* State::max_signal_speed (src/state.rs:121-133) of the current state. */
int32_t cdc_solver_max_signal_speed(cdc_solver_t *solver, double *out);

/*
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the variant of cdc_solver_max_signal_speed that
 * allocates nothing.  `conserved` is a state in device memory, nv doubles per
 * cell ([cell][v], the boundary layout; any cell order), nv = 3 for
 * CDC_ISOTHERMAL, 4 for CDC_EULER; the max signal speed is written to the
 * device double *out.  Work is enqueued on `stream` (a hipStream_t; NULL is
 * the null stream) and the call returns without waiting.  Partial results go
 * to a fixed device array of 1024 doubles (8 KiB) inside the library, one per
 * GPU, kept until the process ends: concurrent calls on one GPU race.
 * ncell must lie in [1, 2^40]; larger or null arguments are rejected.
 */
int32_t cdc_max_signal_speed_noalloc(int32_t kind, double gamma_law_index,
                                     const double *conserved, int64_t ncell,
                                     double *out, void *stream);

/* Tile shape (tx rows along x by ty columns along y, inside one block) of
 * the fused stage kernel.  Supported: 16x32, 32x16, 16x16, 8x32, 32x8 (256
 * threads), 8x16, 16x8 (128 threads), 8x8 (64 threads), and the
 * column-marching strips 32x64, 16x64, 8x64 (64 threads), each dividing
 * block_size; (0, 0) restores the default.  The strips need every array of
 * the grid (conserved planes, aux planes) under 2^32 bytes
 * (max(nv, 3) * N^2 * 8); on larger grids they are rejected with
 * CDC_ERR_INVALID_ARGUMENT and the autotuner skips them.  No reference
 * counterpart: the reference has no tiles.  set_kernel / get_kernel also
 * name the thread count (threads = 0: the shape's only one). */
int32_t cdc_solver_set_tile(cdc_solver_t *solver, int32_t tx, int32_t ty);
int32_t cdc_solver_get_tile(cdc_solver_t *solver, int32_t *tx, int32_t *ty);
int32_t cdc_solver_set_kernel(cdc_solver_t *solver, int32_t tx, int32_t ty, int32_t threads);
int32_t cdc_solver_get_kernel(cdc_solver_t *solver, int32_t *tx, int32_t *ty, int32_t *threads);

/* Autotuner (no reference counterpart): times `reps` stage-kernel launches
 * of every supported tile shape on the current state, writing only scratch
 * buffers (the state is unchanged), and keeps the fastest shape. */
int32_t cdc_solver_autotune(cdc_solver_t *solver, int32_t reps, int32_t *tx,
                            int32_t *ty, double *best_ms);

/* Per-launch timing of the fused stage kernel with HIP events on the
 * handle's stream.  enable=1 clears and starts accumulating. */
int32_t cdc_solver_set_timing(cdc_solver_t *solver, int32_t enable);
int32_t cdc_solver_get_timing(cdc_solver_t *solver, int64_t *launches,
                              double *total_ms, double *cells_per_launch);

/* ---- ensemble: one RK step of many runs per call ------------------------- */
/*
 * No counterpart in the reference (scheme::advance, src/scheme.rs:323-349,
 * steps one run; a parameter study there is one process per run).  An
 * ensemble groups existing solver handles of one hydro kind and rk_order on
 * one device.  cdc_ensemble_advance advances every active member by one RK
 * step at the member's own dt (fold = 1), launching each RK stage once for
 * all members, and leaves each member bitwise where
 * cdc_solver_advance(member, 1, ...) would have.  A member that fails keeps
 * its pre-call state and reports the error cdc_solver_advance would have
 * (codes_out[i], errs_out[i]); the others commit.  Creating the ensemble sets
 * every member's tile shape to one shared shape: the one the members already
 * share when it divides every block size, else the default for the greatest
 * common block size (results of a member depend on its tile shape only
 * through the summation order of the integrated source terms).  Members stay usable on their own; the ensemble does not
 * own them and must be destroyed before them.
 *   active:    n flags (NULL = all), dt_out: n (written for active members),
 *   codes_out: n, errs_out: n (either may be NULL).
 * Returns CDC_OK, or the code of the first failing member in member order.
 */
typedef struct cdc_ensemble cdc_ensemble_t;

int32_t cdc_ensemble_create(cdc_solver_t *const *members, int32_t n,
                            cdc_ensemble_t **out);
void    cdc_ensemble_destroy(cdc_ensemble_t *ensemble);
int32_t cdc_ensemble_advance(cdc_ensemble_t *ensemble, const int32_t *active,
                             double *dt_out, int32_t *codes_out,
                             cdc_error_t *errs_out);
/* Replay each step from a captured HIP graph (default 1) or launch it directly. */
int32_t cdc_ensemble_set_graph(cdc_ensemble_t *ensemble, int32_t enable);
int32_t cdc_ensemble_get_tile(cdc_ensemble_t *ensemble, int32_t *tx, int32_t *ty);
/* Member-steps committed and zone-updates done since creation. */
int32_t cdc_ensemble_get_stats(cdc_ensemble_t *ensemble, int64_t *member_steps,
                               double *zone_updates);

/* ---- learned flux closure for coarse runs -------------------------------- */
/*
 * No counterpart in the reference.  A periodic 3x3 convolutional net
 * (4 -> 32 -> 32 -> 32 -> 6 channels, ReLU; bf16 matrix-core operands, fp32
 * accumulation) maps the state at the start of a coarse step to corrections
 * of the x- and y-face fluxes of (Sigma, Sx, Sy).  A corrected step is the
 * plain RK step followed by
 *   u_v += dt r_v,  r_v = -s_v ((Y_v[i,j] - Y_v[i-1,j]) / dx + (Y_{3+v}[i,j] - Y_{3+v}[i,j-1]) / dy),
 * a flux divergence (mass and momentum conserved), with each face's three
 * corrections scaled down so that a face moves at most kappa min(Sigma_L,
 * Sigma_R) of mass, and that times the larger signal speed |v| + c_s of the
 * two cells of momentum, per step.  Features: a0 ln Sigma,
 * a1 vx, a1 vy, and a2 ln(p / Sigma) (Euler) or a2 ln(-phi) (isothermal: phi the
 * softened two-body potential at the cell centre, -R + (i + 1/2) dx, at the
 * state's time), each clamped to [-16, 16].  Training fits r to
 * (u_truth - u_plain) / dt over stored samples with Adam on the loss
 * (1/P) sum_cells c sum_v w_v (r_v - target_v)^2, c = 1 where the base density
 * is at least mask_density (the cells the correction is applied to), else 0.
 * Parameters in canonical order: per layer W[cout][cin][3][3] then b[cout],
 * layers (4,32), (32,32), (32,32), (32,6) as (cin, cout).
 */
typedef struct cdc_closure cdc_closure_t;

typedef struct cdc_closure_config {
    int32_t  grid;               /* N: the coarse grid is N x N cells */
    int32_t  max_samples;        /* dataset capacity (0: inference only) */
    uint64_t seed;               /* initial weights */
    double   domain_radius;      /* the coarse mesh's (dx = 2R / N) */
    double   feature_scale[3];   /* a0, a1, a2 */
    double   output_scale[3];    /* s_v */
    double   loss_weight[3];     /* w_v */
    double   init_output_gain;   /* multiplies the last layer's initial weights */
    double   limiter;            /* kappa of the face limiter above; 0 = off */
    double   mask_density;       /* faces with min(Sigma_L, Sigma_R) below it are not corrected */
} cdc_closure_config_t;

int32_t cdc_closure_create(const cdc_closure_config_t *config, cdc_closure_t **out);
void    cdc_closure_destroy(cdc_closure_t *closure);
int64_t cdc_closure_num_params(void);
int32_t cdc_closure_get_params(cdc_closure_t *closure, float *params_out);
int32_t cdc_closure_set_params(cdc_closure_t *closure, const float *params);   /* resets Adam */
int32_t cdc_closure_get_config(cdc_closure_t *closure, cdc_closure_config_t *out);
/* Either pointer may be NULL (keep). */
int32_t cdc_closure_set_scales(cdc_closure_t *closure, const double *output_scale,
                               const double *loss_weight);
/* Net input (bf16-rounded, [N*N][4]) and output ([N*N][6]) on a solver's
 * committed state (N x N global grid, i-major). */
int32_t cdc_closure_features(cdc_closure_t *closure, const cdc_solver_t *solver, float *x_out);
int32_t cdc_closure_forward(cdc_closure_t *closure, const cdc_solver_t *solver, float *y_out);
/* Store one training sample: features of base's state, target
 * (truth - plain) / dt for (Sigma, Sx, Sy). */
int32_t cdc_closure_add_sample(cdc_closure_t *closure, const cdc_solver_t *base,
                               const cdc_solver_t *plain, const cdc_solver_t *truth, double dt);
/* count, and the mean square of each target channel over all samples */
int32_t cdc_closure_num_samples(cdc_closure_t *closure, int32_t *count, double *target_ms);
int32_t cdc_closure_clear_samples(cdc_closure_t *closure);
/* Loss and gradient (canonical order) of sample k, no update. */
int32_t cdc_closure_sample_grad(cdc_closure_t *closure, int32_t k, double *loss_out,
                                float *grad_out);
/* `steps` Adam steps (learning rate lr), one sample per step in shuffled
 * epochs; loss_out[steps] (may be NULL). */
int32_t cdc_closure_train(cdc_closure_t *closure, int32_t steps, double lr, uint64_t seed,
                          float *loss_out);
/* One corrected coarse step of solver (cdc_solver_advance(solver, 1) plus
 * the correction).  Errors as cdc_solver_advance; with an all-zero net the
 * result is bitwise that of cdc_solver_advance. */
int32_t cdc_closure_advance(cdc_closure_t *closure, cdc_solver_t *solver, double *dt_out,
                            cdc_error_t *err);
/* `fold` corrected coarse steps at one dt (cdc_solver_advance(solver, fold)
 * with the correction after every step, the net evaluated on each step's
 * starting state), as one captured graph replayed per call and one host
 * synchronisation; the next call's dt is the max signal speed of the
 * corrected result.  Transactional and with the errors of
 * cdc_solver_advance; with an all-zero net the result is bitwise that of
 * cdc_solver_advance(solver, fold).  cdc_closure_advance is fold 1. */
int32_t cdc_closure_advance_fold(cdc_closure_t *closure, cdc_solver_t *solver, int32_t fold,
                                 double *dt_out, cdc_error_t *err);
/* Correct every `interval`-th step only (default 1): the steps of a call go
 * in groups of `interval` (the last group of a call may be shorter); the
 * net is evaluated on a group's starting state and its limited correction
 * applied to the group's plain result over the group's time (steps x dt).
 * No counterpart in the reference (the learned closure is an extension);
 * for callers that trade the correction's time resolution for wall time. */
int32_t cdc_closure_set_interval(cdc_closure_t *closure, int32_t interval);

/* ---- host utilities (no GPU) -------------------------------------------- */

/*
 * Re-grid a state in the boundary layout between two meshes of one domain
 * (the same rules as cdc_solver_resample_from): u_src [nbS^2][bsS][bsS][nv]
 * -> u_dst [nbD^2][bsD][bsD][nv].  Per-block ist [.][7][nv] and oec [.][7][4]
 * are copied when num_blocks is kept (the up-sampler's rule), summed over
 * merged blocks when each coarse block covers whole fine blocks, and zeroed
 * otherwise; the ist/oec pointers may be NULL.
 */
int32_t cdc_resample_state(const cdc_mesh_t *src, const cdc_mesh_t *dst, int32_t nv,
                           const double *u_src, const double *ist_src,
                           const double *oec_src, double *u_dst, double *ist_dst,
                           double *oec_dst);

/* Physics::orbital_state_from_time (src/physics.rs:419-421). */
int32_t cdc_orbital_state_from_time(const cdc_physics_t *physics, double time,
                                    cdc_point_mass_t out[2]);

/*
 * Seeded cavity disk: an initial model with no reference counterpart, an
 * InitialModel (src/traits.rs:105-116) written out per cell.
 *   Sigma(r) = max(floor, exp(-(r_cav / r)^xi)) * (1 + amp * n_sigma)
 *   v_phi    = sqrt(M r^2 / (r^2 + r_s^2)^(3/2))   (about the centre of mass)
 *   p        = Sigma / sqrt(r^2 + r_s^2) / (gamma M_a^2) * (1 + amp * n_p)   (Euler:
 *              c_s^2 = -phi / M_a^2 of the softened unit mass, the isothermal law)
 * and, when taper_radius > 0, v_phi multiplied by 1/(1 + exp((r - r_t)/w)) so
 * that the flow vanishes at the periodic edges (inside the buffer zone),
 * with n_sigma, n_p uniform in [-1, 1) drawn by a counter-based generator
 * (SplitMix64) from (seed, stream, global cell index), so the state is a pure
 * function of the seed and the global grid: block-decomposition invariant and
 * reproducible on restart.
 */
typedef struct cdc_cavity_model {
    double   cavity_radius;      /* r_cav, 2.5 */
    double   cavity_exponent;    /* xi, 12 */
    double   density_floor;      /* 1e-4 */
    double   noise_amplitude;    /* 1e-3 */
    double   softening_length;   /* r_s in v_phi */
    double   mach_number;        /* M_a for the Euler pressure */
    double   taper_radius;       /* v_phi *= 1/(1 + exp((r - r_t)/w)); 0 = no taper */
    double   taper_width;        /* w */
    uint64_t seed;
} cdc_cavity_model_t;

int32_t cdc_model_seeded_cavity(const cdc_mesh_t *mesh, const cdc_hydro_t *hydro,
                                const cdc_cavity_model_t *model,
                                double *conserved_out);

#ifdef __cplusplus
}
#endif

#endif /* CDC_SOLVER_H */
