/*
 * cdc_oracle.c -- CPU restatement of the reference's per-block Runge-Kutta
 * update (clemson-cal/app-binary2d: src/scheme.rs, src/physics.rs,
 * src/state.rs, src/mesh.rs).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity reference for the HIP path
 * and the timed CPU baseline.  Nothing in the product links, loads or calls
 * it; only the tests, the smoke check and the benchmark's CPU-baseline leg do.
 *
 * Structure follows the reference one to one:
 *   oracle_advance          scheme::advance              scheme.rs:323-349
 *   rk_step                 RungeKuttaOrder driver       (godunov-core, see below)
 *   try_advance_rk          try_advance_rk               scheme.rs:88-173
 *   block_primitive         try_block_primitive          scheme.rs:185-199
 *   block_fluxes            compute_block_fluxes         scheme.rs:201-271
 *   block_update            compute_block_updated_solution scheme.rs:273-316
 *   weighted_average        WeightedAverage for State    state.rs:183-219
 *   max_signal_speed        State::max_signal_speed      state.rs:121-133
 * Arithmetic is IEEE binary64 with no contraction (built -ffp-contract=off),
 * in the operation order the reference source spells out.
 *
 * The reference's arithmetic crates are git dependencies that are not in the
 * reference tree (Cargo.toml:29-33): godunov-core @d8dd524, hydro-iso2d
 * @eb18a38, hydro-euler @fc929f9, kepler-two-body @2d8a0f5, ndarray-ops
 * @725bc21.  Their published algorithms are restated here:
 *   plm_gradient   generalized minmod(theta(b-a), (c-a)/2, theta(c-b))
 *   riemann_hlle   HLLE (Harten-Lax-van Leer-Einfeldt, Toro ch. 10) with
 *                  outer wave speeds v_n -/+ c_s
 *   RK1/2/3        SSP Runge-Kutta as weighted averages b = 1/2 (RK2) and
 *                  b = 3/4, 1/3 (RK3) of the stage state with the step base
 *   Kepler orbit   Newton solve of E - e sin E = n t, pericentre on +x at
 *                  t = 0, body 1 at -(m2/M) r_rel, body 2 at +(m1/M) r_rel,
 *                  m1 = M/(1+q), m2 = M q/(1+q)
 *   Plummer        phi = -m / sqrt(r^2 + s^2), a = -m d / (t sqrt(t)), t = r^2+s^2
 *   perturb        add (dm, dp) to each body at the orbital state of time t,
 *                  recover (a, M, q, e) from the relative orbit (energy ->
 *                  a, eccentricity vector -> e); unbound if E >= 0
 *   linspace       x_k = a + k (b - a)/(n - 1) (ndarray 0.14)
 *   adjacent_mean  (x_k + x_{k+1}) / 2
 * What pins each of them is listed in the tests (tests/test_oracle_pins.py).
 * Parity unpinned (only the absent crates could pin it; the HIP path shares
 * the convention, so GPU parity is unaffected): the order of the additions
 * inside riemann_hlle and plm_gradient (the pins fix their values to
 * rounding), and the Kepler phase at t = 0 and the body order above.
 * The Euler primitive recovery p = (E - Sigma |v|^2 / 2)(gamma - 1) and the
 * Euler max signal speed max(|vx|, |vy|) + c_s are chosen conventions whose
 * exact operation order only the absent crates could pin.
 */
#include "../include/cdc_solver.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NITEM 7
#define ORBITAL_PERIOD (2.0 * 3.14159265358979323846)

typedef struct { double v[4]; } vec4;        /* Conserved or Primitive */
typedef struct { double m, x, y, vx, vy; } pmass;
typedef struct { pmass p[2]; } orbit_state;   /* kepler_two_body::OrbitalState */
typedef struct { vec4 it[NITEM]; } itemized;  /* ItemizedChange<Conserved> */
typedef struct { double e[4]; } elements;     /* OrbitalElements (a, M, q, e) */

typedef struct {
    int code;            /* CDC_ERR_* */
    double value, x, y;
} herr;

typedef struct {
    double time;
    int64_t num, den;    /* Rational64 iteration */
    double *u;           /* [nblk][bs][bs][nv] */
    double *ist;         /* [nblk][7][nv] */
    double *oec;         /* [nblk][7][4] */
    double *tq;          /* [nblk][2] torque impulse on body 1, 2 (extension) */
} ostate;

typedef struct oracle {
    cdc_mesh_t mesh;
    cdc_hydro_t hydro;
    cdc_physics_t phys;
    int nv, nb, bs, nblk;
    size_t ncell;        /* per block bs*bs */
    double *initial;     /* BlockData::initial_conserved, all blocks */
    ostate st;
    ostate work[3];
    double *prim;        /* [nblk][bs][bs][nv] */
    herr *perr;          /* per block, P-phase */
    herr *uerr;          /* per block, U-phase */
} oracle_t;

/* ------------------------------------------------------------------------- */
/* Rational64 helpers (num::rational, src/state.rs:215)                      */

static int64_t gcd64(int64_t a, int64_t b) {
    if (a < 0) a = -a;
    if (b < 0) b = -b;
    while (b) { int64_t t = a % b; a = b; b = t; }
    return a ? a : 1;
}

static void rat_norm(int64_t *n, int64_t *d) {
    int64_t g = gcd64(*n, *d);
    *n /= g; *d /= g;
    if (*d < 0) { *n = -*n; *d = -*d; }
}

/* iteration * (1 - br) + s0.iteration * br, br = bn/bd */
static void rat_wavg(int64_t n1, int64_t d1, int64_t n0, int64_t d0,
                     int64_t bn, int64_t bd, int64_t *on, int64_t *od) {
    int64_t an = (bd - bn) * n1, ad = bd * d1;  /* s1 * (1 - b) */
    int64_t cn = bn * n0, cd = bd * d0;         /* s0 * b */
    rat_norm(&an, &ad);
    rat_norm(&cn, &cd);
    int64_t rn = an * cd + cn * ad, rd = ad * cd;
    rat_norm(&rn, &rd);
    *on = rn; *od = rd;
}

/* ------------------------------------------------------------------------- */
/* kepler-two-body restatement                                               */

void oracle_orbital_elements(const cdc_physics_t *ph, double out[4]) {
    /* Physics::orbital_elements, src/physics.rs:411-417 */
    if (ph->one_body) {
        out[0] = 1e-12; out[1] = 1.0; out[2] = 1.0; out[3] = 0.0;
    } else {
        out[0] = 1.0; out[1] = 1.0; out[2] = ph->binary_mass_ratio; out[3] = ph->binary_eccentricity;
    }
}

static void orbital_state_from_elements(const double el[4], double t, orbit_state *s) {
    double a = el[0], M = el[1], q = el[2], e = el[3];
    double m1 = M / (1.0 + q);
    double m2 = M * q / (1.0 + q);
    double n = sqrt(M / (a * a * a));
    double mean = n * t;
    double E = mean;
    for (int k = 0; k < 64; ++k) {
        double f = E - e * sin(E) - mean;
        double fp = 1.0 - e * cos(E);
        double dE = f / fp;
        E = E - dE;
        if (fabs(dE) <= 1e-15 * (1.0 + fabs(E))) break;
    }
    double cE = cos(E), sE = sin(E);
    double b = sqrt(1.0 - e * e);
    double x = a * (cE - e);
    double y = a * b * sE;
    double den = 1.0 - e * cE;
    double vx = -a * n * sE / den;
    double vy = a * n * b * cE / den;
    double f1 = m2 / M, f2 = m1 / M;
    s->p[0].m = m1; s->p[0].x = -f1 * x; s->p[0].y = -f1 * y; s->p[0].vx = -f1 * vx; s->p[0].vy = -f1 * vy;
    s->p[1].m = m2; s->p[1].x =  f2 * x; s->p[1].y =  f2 * y; s->p[1].vx =  f2 * vx; s->p[1].vy =  f2 * vy;
}

static void orbital_state_from_time(const cdc_physics_t *ph, double t, orbit_state *s) {
    double el[4];
    oracle_orbital_elements(ph, el);
    orbital_state_from_elements(el, t, s);
}

void oracle_orbital_state_from_time(const cdc_physics_t *ph, double t, double out[10]) {
    orbit_state s;
    orbital_state_from_time(ph, t, &s);
    memcpy(out, &s, sizeof(s));
}

static double softening_length(const cdc_physics_t *ph) {
    /* src/physics.rs:399-405 */
    return ph->has_softening_length ? ph->softening_length : ph->sink_radius;
}

static double pm_potential(const pmass *p, double x, double y, double rs) {
    double dx = x - p->x, dy = y - p->y;
    double r2 = dx * dx + dy * dy;
    return -p->m / sqrt(r2 + rs * rs);
}

static void pm_acceleration(const pmass *p, double x, double y, double rs, double *ax, double *ay) {
    double dx = x - p->x, dy = y - p->y;
    double r2 = dx * dx + dy * dy;
    double t = r2 + rs * rs;
    double mag = -p->m / (t * sqrt(t));
    *ax = mag * dx;
    *ay = mag * dy;
}

static double gravitational_potential(const orbit_state *s, double x, double y, double rs) {
    return pm_potential(&s->p[0], x, y, rs) + pm_potential(&s->p[1], x, y, rs);
}

void oracle_gravity(const double state[10], double x, double y, double rs,
                    double *phi, double acc[4]) {
    orbit_state s;
    memcpy(&s, state, sizeof(s));
    *phi = gravitational_potential(&s, x, y, rs);
    pm_acceleration(&s.p[0], x, y, rs, &acc[0], &acc[1]);
    pm_acceleration(&s.p[1], x, y, rs, &acc[2], &acc[3]);
}

/* OrbitalElements::perturb at the orbital state s of time t.  Returns 0, or
 * CDC_ERR_UNBOUND_ORBIT when the perturbed binary is unbound. */
static int perturb(const orbit_state *s, double dm1, double dm2, double dpx1,
                   double dpx2, double dpy1, double dpy2, double out[4]) {
    double m1 = s->p[0].m + dm1;
    double m2 = s->p[1].m + dm2;
    double vx1 = (s->p[0].m * s->p[0].vx + dpx1) / m1;
    double vy1 = (s->p[0].m * s->p[0].vy + dpy1) / m1;
    double vx2 = (s->p[1].m * s->p[1].vx + dpx2) / m2;
    double vy2 = (s->p[1].m * s->p[1].vy + dpy2) / m2;
    double M = m1 + m2;
    double q = m2 / m1;
    double rx = s->p[1].x - s->p[0].x;
    double ry = s->p[1].y - s->p[0].y;
    double vx = vx2 - vx1;
    double vy = vy2 - vy1;
    double r = sqrt(rx * rx + ry * ry);
    double energy = 0.5 * (vx * vx + vy * vy) - M / r;
    if (!(energy < 0.0) || !(m1 > 0.0) || !(m2 > 0.0)) return CDC_ERR_UNBOUND_ORBIT;
    double a = -M / (2.0 * energy);
    double h = rx * vy - ry * vx;
    double ex = vy * h / M - rx / r;
    double ey = -vx * h / M - ry / r;
    out[0] = a;
    out[1] = M;
    out[2] = q;
    out[3] = sqrt(ex * ex + ey * ey);
    return 0;
}

int oracle_perturb(const double state[10], const double el[4], const double delta[6], double out[4]) {
    orbit_state s;
    double p[4];
    memcpy(&s, state, sizeof(s));
    int rc = perturb(&s, delta[0], delta[1], delta[2], delta[3], delta[4], delta[5], p);
    if (rc) return rc;
    for (int k = 0; k < 4; ++k) out[k] = p[k] - el[k];
    return 0;
}

/* ItemizedChange::perturbation, src/state.rs:258-282 */
static int perturbation(const itemized *ds, const orbit_state *s, const double el[4], elements de[NITEM]) {
    double p[4];
    memset(de, 0, sizeof(elements) * NITEM);
    const int items[4] = {CDC_ITEM_SINK1, CDC_ITEM_SINK2, CDC_ITEM_GRAV1, CDC_ITEM_GRAV2};
    for (int k = 0; k < 4; ++k) {
        const vec4 *d = &ds->it[items[k]];
        int rc;
        if (k % 2 == 0) /* pert1: body 1 receives (-dm, -dp) */
            rc = perturb(s, -d->v[0], 0.0, -d->v[1], 0.0, -d->v[2], 0.0, p);
        else            /* pert2 */
            rc = perturb(s, 0.0, -d->v[0], 0.0, -d->v[1], 0.0, -d->v[2], p);
        if (rc) return rc;
        for (int c = 0; c < 4; ++c) de[items[k]].e[c] = p[c] - el[c];
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* hydro-iso2d / hydro-euler restatement                                     */

static vec4 iso_to_primitive(vec4 u) {
    vec4 p = {{u.v[0], u.v[1] / u.v[0], u.v[2] / u.v[0], 0.0}};
    return p;
}

static vec4 iso_to_conserved(vec4 p) {
    vec4 u = {{p.v[0], p.v[0] * p.v[1], p.v[0] * p.v[2], 0.0}};
    return u;
}

static vec4 euler_to_primitive(vec4 u, double gm) {
    double d = u.v[0];
    double vx = u.v[1] / d;
    double vy = u.v[2] / d;
    double p = (u.v[3] - 0.5 * d * (vx * vx + vy * vy)) * (gm - 1.0);
    vec4 r = {{d, vx, vy, p}};
    return r;
}

static vec4 euler_to_conserved(vec4 p, double gm) {
    double d = p.v[0], vx = p.v[1], vy = p.v[2];
    vec4 u = {{d, d * vx, d * vy, 0.5 * d * (vx * vx + vy * vy) + p.v[3] / (gm - 1.0)}};
    return u;
}

static double euler_specific_internal_energy(vec4 p, double gm) {
    return p.v[3] / (p.v[0] * (gm - 1.0));
}

static double euler_specific_kinetic_energy(vec4 p) {
    return 0.5 * (p.v[1] * p.v[1] + p.v[2] * p.v[2]);
}

/* HLLE, isothermal: p = Sigma cs2, waves v_n -/+ cs */
static vec4 iso_riemann_hlle(vec4 pl, vec4 pr, int axis, double cs2) {
    double cs = sqrt(cs2);
    vec4 ul = iso_to_conserved(pl), ur = iso_to_conserved(pr);
    double vnl = pl.v[1 + axis], vnr = pr.v[1 + axis];
    double ppl = pl.v[0] * cs2, ppr = pr.v[0] * cs2;
    vec4 fl = {{ul.v[0] * vnl, ul.v[1] * vnl, ul.v[2] * vnl, 0.0}};
    vec4 fr = {{ur.v[0] * vnr, ur.v[1] * vnr, ur.v[2] * vnr, 0.0}};
    fl.v[1 + axis] += ppl;
    fr.v[1 + axis] += ppr;
    double alm = vnl - cs, alp = vnl + cs;
    double arm = vnr - cs, arp = vnr + cs;
    double ap = fmax(fmax(alp, arp), 0.0);
    double am = fmin(fmin(alm, arm), 0.0);
    vec4 f;
    for (int k = 0; k < 3; ++k)
        f.v[k] = (fl.v[k] * ap - fr.v[k] * am - (ul.v[k] - ur.v[k]) * ap * am) / (ap - am);
    f.v[3] = 0.0;
    return f;
}

/* HLLE, gamma-law: waves v_n -/+ c_s, c_s^2 = gamma p / Sigma */
static vec4 euler_riemann_hlle(vec4 pl, vec4 pr, int axis, double gm) {
    vec4 ul = euler_to_conserved(pl, gm), ur = euler_to_conserved(pr, gm);
    double vnl = pl.v[1 + axis], vnr = pr.v[1 + axis];
    double csl = sqrt(gm * pl.v[3] / pl.v[0]);
    double csr = sqrt(gm * pr.v[3] / pr.v[0]);
    vec4 fl = {{ul.v[0] * vnl, ul.v[1] * vnl, ul.v[2] * vnl, (ul.v[3] + pl.v[3]) * vnl}};
    vec4 fr = {{ur.v[0] * vnr, ur.v[1] * vnr, ur.v[2] * vnr, (ur.v[3] + pr.v[3]) * vnr}};
    fl.v[1 + axis] += pl.v[3];
    fr.v[1 + axis] += pr.v[3];
    double alm = vnl - csl, alp = vnl + csl;
    double arm = vnr - csr, arp = vnr + csr;
    double ap = fmax(fmax(alp, arp), 0.0);
    double am = fmin(fmin(alm, arm), 0.0);
    vec4 f;
    for (int k = 0; k < 4; ++k)
        f.v[k] = (fl.v[k] * ap - fr.v[k] * am - (ul.v[k] - ur.v[k]) * ap * am) / (ap - am);
    return f;
}

void oracle_riemann_hlle(int kind, const double pl[4], const double pr[4], int axis,
                         double cs2_or_gamma, double out[4]) {
    vec4 a, b, f;
    memcpy(a.v, pl, sizeof(a.v));
    memcpy(b.v, pr, sizeof(b.v));
    f = kind == CDC_EULER ? euler_riemann_hlle(a, b, axis, cs2_or_gamma)
                          : iso_riemann_hlle(a, b, axis, cs2_or_gamma);
    memcpy(out, f.v, sizeof(f.v));
}

/* godunov_core::piecewise_linear generalized minmod */
static double minmod3(double a, double b, double c) {
    if (a > 0.0 && b > 0.0 && c > 0.0) return fmin(a, fmin(b, c));
    if (a < 0.0 && b < 0.0 && c < 0.0) return fmax(a, fmax(b, c));
    return 0.0;
}

double oracle_plm_gradient_scalar(double theta, double a, double b, double c) {
    return minmod3(theta * (b - a), 0.5 * (c - a), theta * (c - b));
}

static vec4 plm_gradient(double theta, const vec4 *a, const vec4 *b, const vec4 *c, int nv) {
    vec4 g = {{0, 0, 0, 0}};
    for (int k = 0; k < nv; ++k) g.v[k] = oracle_plm_gradient_scalar(theta, a->v[k], b->v[k], c->v[k]);
    return g;
}

/* ------------------------------------------------------------------------- */
/* Physics (src/physics.rs)                                                  */

/* CellData::stress_field, src/physics.rs:322-340.  row/col: 0 = X, 1 = Y */
static double stress_field(const vec4 *pc, const vec4 *gx, const vec4 *gy, double nu,
                           double lambda, double dx, double dy, int row, int col) {
    double shear, bulk;
    if (row == 0 && col == 0) {
        shear = 4.0 / 3.0 * gx->v[1] / dx - 2.0 / 3.0 * gy->v[2] / dy;
        bulk = gx->v[1] / dx + gy->v[2] / dy;
    } else if (row == 1 && col == 1) {
        shear = -2.0 / 3.0 * gx->v[1] / dx + 4.0 / 3.0 * gy->v[2] / dy;
        bulk = gx->v[1] / dx + gy->v[2] / dy;
    } else {
        shear = 1.0 / 1.0 * gx->v[2] / dx + 1.0 / 1.0 * gy->v[1] / dy;
        bulk = 0.0;
    }
    return pc->v[0] * (nu * shear + lambda * bulk);
}

double oracle_stress_field(const double pc[4], const double gx[4], const double gy[4], double nu,
                           double lambda, double dx, double dy, int row, int col) {
    vec4 a, b, c;
    memcpy(a.v, pc, sizeof(a.v));
    memcpy(b.v, gx, sizeof(b.v));
    memcpy(c.v, gy, sizeof(c.v));
    return stress_field(&a, &b, &c, nu, lambda, dx, dy, row, col);
}

/* Physics::sink_kernel, src/physics.rs:388-397 */
static double sink_kernel(const cdc_physics_t *ph, double dx, double dy) {
    double r2 = dx * dx + dy * dy;
    double s2 = ph->sink_radius * ph->sink_radius;
    if (r2 < s2) {
        double w = 1.0 - r2 / s2;
        return ph->sink_rate * (w * w);
    }
    return 0.0;
}

double oracle_sink_kernel(const cdc_physics_t *ph, double dx, double dy) {
    return sink_kernel(ph, dx, dy);
}

typedef struct { double fx1, fy1, fx2, fy2, sink_rate1, sink_rate2, buffer_rate; } source_terms_t;

/* Physics::source_terms, src/physics.rs:423-457 */
static source_terms_t physics_source_terms(const cdc_physics_t *ph, const cdc_mesh_t *mesh,
                                           const orbit_state *s, double x, double y, double sd) {
    source_terms_t st;
    double rs = softening_length(ph);
    double ax1, ay1, ax2, ay2;
    pm_acceleration(&s->p[0], x, y, rs, &ax1, &ay1);
    pm_acceleration(&s->p[1], x, y, rs, &ax2, &ay2);
    st.fx1 = sd * ax1;
    st.fy1 = sd * ay1;
    st.fx2 = sd * ax2;
    st.fy2 = sd * ay2;
    st.sink_rate1 = sink_kernel(ph, x - s->p[0].x, y - s->p[0].y);
    st.sink_rate2 = sink_kernel(ph, x - s->p[1].x, y - s->p[1].y);
    double r = sqrt(x * x + y * y);
    double yb = (r - mesh->domain_radius) / ph->buffer_scale;
    double R = mesh->domain_radius;
    double omega_outer = sqrt((s->p[0].m + s->p[1].m) / (R * R * R));
    st.buffer_rate = 0.5 * ph->buffer_rate * (1.0 + tanh(yb)) * omega_outer;
    return st;
}

static vec4 v_scale(vec4 a, double s, int nv) {
    for (int k = 0; k < nv; ++k) a.v[k] = a.v[k] * s;
    return a;
}

static vec4 v_add(vec4 a, vec4 b, int nv) {
    for (int k = 0; k < nv; ++k) a.v[k] = a.v[k] + b.v[k];
    return a;
}

static vec4 v_sub(vec4 a, vec4 b, int nv) {
    for (int k = 0; k < nv; ++k) a.v[k] = a.v[k] - b.v[k];
    return a;
}

/* Isothermal::source_terms, src/physics.rs:528-582 */
static itemized iso_source_terms(const cdc_physics_t *ph, const cdc_mesh_t *mesh, vec4 u, vec4 u0,
                                 double x, double y, double dt, const orbit_state *s) {
    itemized r;
    vec4 zero = {{0, 0, 0, 0}};
    vec4 fake_du;
    if (u.v[0] < u0.v[0] * ph->fake_mass_threshold)
        fake_du = zero;
    else
        fake_du = v_scale(v_scale(u0, ph->fake_mass_rate, 3), dt, 3);
    source_terms_t st = physics_source_terms(ph, mesh, s, x, y, u.v[0]);
    vec4 sink1, sink2;
    if (ph->sink_model == CDC_SINK_IMPARTIAL) {
        sink1 = v_scale(u, -st.sink_rate1, 3);
        sink2 = v_scale(u, -st.sink_rate2, 3);
    } else {
        vec4 p = iso_to_primitive(u);
        double vx = p.v[1], vy = p.v[2];
        double x1 = s->p[0].x, y1 = s->p[0].y, x2 = s->p[1].x, y2 = s->p[1].y;
        double vx1 = ph->one_body ? 0.0 : s->p[0].vx, vy1 = ph->one_body ? 0.0 : s->p[0].vy;
        double vx2 = ph->one_body ? 0.0 : s->p[1].vx, vy2 = ph->one_body ? 0.0 : s->p[1].vy;
        double r1 = sqrt((x - x1) * (x - x1) + (y - y1) * (y - y1));
        double r2 = sqrt((x - x2) * (x - x2) + (y - y2) * (y - y2));
        double r1x = (x - x1) / r1, r1y = (y - y1) / r1;
        double r2x = (x - x2) / r2, r2y = (y - y2) / r2;
        double dvdotr1 = (vx - vx1) * r1x + (vy - vy1) * r1y;
        double dvdotr2 = (vx - vx2) * r2x + (vy - vy2) * r2y;
        double vx1c = dvdotr1 * r1x + vx1, vy1c = dvdotr1 * r1y + vy1;
        double vx2c = dvdotr2 * r2x + vx2, vy2c = dvdotr2 * r2y + vy2;
        double sd = p.v[0];
        vec4 s1v = {{sd, sd * vx1c, sd * vy1c, 0.0}};
        vec4 s2v = {{sd, sd * vx2c, sd * vy2c, 0.0}};
        sink1 = v_scale(s1v, -st.sink_rate1, 3);
        sink2 = v_scale(s2v, -st.sink_rate2, 3);
    }
    vec4 g1 = {{0.0, st.fx1, st.fy1, 0.0}};
    vec4 g2 = {{0.0, st.fx2, st.fy2, 0.0}};
    r.it[CDC_ITEM_GRAV1] = v_scale(g1, dt, 3);
    r.it[CDC_ITEM_GRAV2] = v_scale(g2, dt, 3);
    r.it[CDC_ITEM_SINK1] = v_scale(sink1, dt, 3);
    r.it[CDC_ITEM_SINK2] = v_scale(sink2, dt, 3);
    r.it[CDC_ITEM_BUFFER] = v_scale(v_sub(u, u0, 3), -dt * st.buffer_rate, 3);
    r.it[CDC_ITEM_COOLING] = zero;
    r.it[CDC_ITEM_FAKE_MASS] = fake_du;
    return r;
}

/* Euler::source_terms, src/physics.rs:683-774 (+ beta cooling, an extension) */
static itemized euler_source_terms(const cdc_hydro_t *hy, const cdc_physics_t *ph, const cdc_mesh_t *mesh,
                                   vec4 u0, vec4 bg, double x, double y, double t, double dt,
                                   const orbit_state *s) {
    itemized r;
    vec4 zero = {{0, 0, 0, 0}};
    double gm = hy->gamma_law_index;
    source_terms_t st = physics_source_terms(ph, mesh, s, x, y, u0.v[0]);
    vec4 p0 = euler_to_primitive(u0, gm);
    double vx = p0.v[1];
    double vy = p0.v[2];
    double e0 = euler_specific_internal_energy(p0, gm);
    double ec;
    if (hy->cooling_model == CDC_COOLING_BETA) {
        double rs = softening_length(ph);
        double r2 = x * x + y * y + rs * rs;
        double omega = sqrt((s->p[0].m + s->p[1].m) / (r2 * sqrt(r2)));
        double et = euler_specific_internal_energy(euler_to_primitive(bg, gm), gm);
        ec = et + (e0 - et) * exp(-dt * omega / hy->cooling_beta);
    } else {
        double f = hy->cooling_slow_start ? exp(-ORBITAL_PERIOD / (t + dt)) : 1.0;
        double sdi = hy->density_index == 1 ? p0.v[0] : p0.v[0] * p0.v[0];
        if (hy->density_index != 1 && hy->density_index != 2) sdi = pow(p0.v[0], (double)hy->density_index);
        double a = hy->cooling_coefficient / sdi * f;
        ec = e0 * pow(1.0 + 3.0 * a * (e0 * e0 * e0) * dt, -1.0 / 3.0);
    }
    if (ph->has_mach_ceiling) {
        double ek = euler_specific_kinetic_energy(p0);
        double mc = ph->mach_ceiling;
        ec = fmax(ec, 2.0 * ek / gm / (gm - 1.0) / (mc * mc));
    }
    vec4 pc = {{p0.v[0], p0.v[1], p0.v[2], p0.v[0] * ec * (gm - 1.0)}};
    vec4 uc = euler_to_conserved(pc, gm);
    vec4 sink1, sink2;
    if (ph->sink_model == CDC_SINK_IMPARTIAL) {
        sink1 = v_scale(u0, -st.sink_rate1, 4);
        sink2 = v_scale(u0, -st.sink_rate2, 4);
    } else {
        double x1 = s->p[0].x, y1 = s->p[0].y, x2 = s->p[1].x, y2 = s->p[1].y;
        double vx1 = ph->one_body ? 0.0 : s->p[0].vx, vy1 = ph->one_body ? 0.0 : s->p[0].vy;
        double vx2 = ph->one_body ? 0.0 : s->p[1].vx, vy2 = ph->one_body ? 0.0 : s->p[1].vy;
        double r1 = sqrt((x - x1) * (x - x1) + (y - y1) * (y - y1));
        double r2 = sqrt((x - x2) * (x - x2) + (y - y2) * (y - y2));
        double r1x = (x - x1) / r1, r1y = (y - y1) / r1;
        double r2x = (x - x2) / r2, r2y = (y - y2) / r2;
        double dvdotr1 = (vx - vx1) * r1x + (vy - vy1) * r1y;
        double dvdotr2 = (vx - vx2) * r2x + (vy - vy2) * r2y;
        double vx1c = dvdotr1 * r1x + vx1, vy1c = dvdotr1 * r1y + vy1;
        double vx2c = dvdotr2 * r2x + vx2, vy2c = dvdotr2 * r2y + vy2;
        double v1csq = vx1c * vx1c + vy1c * vy1c;
        double v2csq = vx2c * vx2c + vy2c * vy2c;
        double sd = p0.v[0];
        vec4 s1v = {{sd, sd * vx1c, sd * vy1c, sd * e0 + 0.5 * sd * v1csq}};
        vec4 s2v = {{sd, sd * vx2c, sd * vy2c, sd * e0 + 0.5 * sd * v2csq}};
        sink1 = v_scale(s1v, -st.sink_rate1, 4);
        sink2 = v_scale(s2v, -st.sink_rate2, 4);
    }
    vec4 g1 = {{0.0, st.fx1, st.fy1, st.fx1 * vx + st.fy1 * vy}};
    vec4 g2 = {{0.0, st.fx2, st.fy2, st.fx2 * vx + st.fy2 * vy}};
    r.it[CDC_ITEM_GRAV1] = v_scale(g1, dt, 4);
    r.it[CDC_ITEM_GRAV2] = v_scale(g2, dt, 4);
    r.it[CDC_ITEM_SINK1] = v_scale(sink1, dt, 4);
    r.it[CDC_ITEM_SINK2] = v_scale(sink2, dt, 4);
    r.it[CDC_ITEM_BUFFER] = v_scale(v_sub(u0, bg, 4), -dt * st.buffer_rate, 4);
    r.it[CDC_ITEM_COOLING] = v_sub(uc, u0, 4);
    r.it[CDC_ITEM_FAKE_MASS] = zero;
    return r;
}

/* Per-cell source terms for tests: out[7][4] */
void oracle_source_terms(const cdc_mesh_t *mesh, const cdc_hydro_t *hy, const cdc_physics_t *ph,
                         const double u[4], const double u0[4], double x, double y, double t,
                         double dt, const double state[10], double out[28]) {
    orbit_state s;
    vec4 a, b;
    itemized r;
    memcpy(&s, state, sizeof(s));
    memcpy(a.v, u, sizeof(a.v));
    memcpy(b.v, u0, sizeof(b.v));
    if (hy->kind == CDC_EULER)
        r = euler_source_terms(hy, ph, mesh, a, b, x, y, t, dt, &s);
    else
        r = iso_source_terms(ph, mesh, a, b, x, y, dt, &s);
    memcpy(out, &r, sizeof(r));
}

/* ------------------------------------------------------------------------- */
/* Mesh geometry (src/mesh.rs:45-92)                                         */

static double block_length(const cdc_mesh_t *m) { return 2.0 * m->domain_radius / (double)m->num_blocks; }
static double cell_spacing(const cdc_mesh_t *m) { return block_length(m) / (double)m->block_size; }

/* Vertices of block index `bi` along one axis: linspace(start, start+L, bs+1) */
static void block_vertices(const cdc_mesh_t *m, int bi, double *xv) {
    double L = block_length(m);
    double a = -m->domain_radius + (double)bi * L;
    double b = a + L;
    int n = m->block_size + 1;
    double step = (b - a) / (double)(n - 1);
    for (int k = 0; k < n; ++k) xv[k] = a + step * (double)k;
}

void oracle_block_vertices(const cdc_mesh_t *m, int bi, double *xv) { block_vertices(m, bi, xv); }

/* ------------------------------------------------------------------------- */
/* Hydrodynamics trait dispatch                                              */

static int try_to_primitive(const cdc_hydro_t *hy, vec4 u, vec4 *p, herr *e) {
    if (hy->kind == CDC_ISOTHERMAL) {
        /* src/physics.rs:489-503 */
        if (hy->has_density_floor && u.v[0] < hy->density_floor) {
            u.v[1] = 0.0;
            u.v[2] = 0.0;
            u.v[0] = hy->density_floor;
        }
        if (u.v[0] < 0.0) {
            e->code = CDC_ERR_NEGATIVE_DENSITY;
            e->value = u.v[0];
            return 1;
        }
        *p = iso_to_primitive(u);
        return 0;
    }
    /* src/physics.rs:631-660 */
    double gm = hy->gamma_law_index;
    vec4 q = euler_to_primitive(u, gm);
    if (hy->has_density_floor && q.v[0] < hy->density_floor) {
        q.v[0] = hy->density_floor;
        q.v[1] = 0.0;
        q.v[2] = 0.0;
        q.v[3] = pow(hy->density_floor, gm);
    }
    if (q.v[0] < 0.0) {
        e->code = CDC_ERR_NEGATIVE_DENSITY;
        e->value = q.v[0];
        return 1;
    }
    if (hy->has_pressure_floor && q.v[3] < hy->pressure_floor) q.v[3] = hy->pressure_floor;
    if (q.v[3] <= 0.0) {
        e->code = CDC_ERR_NEGATIVE_PRESSURE;
        e->value = q.v[3];
        return 1;
    }
    *p = q;
    return 0;
}

static double hydro_max_signal_speed(const cdc_hydro_t *hy, vec4 u) {
    if (hy->kind == CDC_ISOTHERMAL) {
        /* src/physics.rs:521-526 */
        vec4 p = iso_to_primitive(u);
        return fmax(fabs(p.v[1]), fabs(p.v[2]));
    }
    double gm = hy->gamma_law_index;
    vec4 p = euler_to_primitive(u, gm);
    double cs = sqrt(gm * p.v[3] / p.v[0]);
    return fmax(fabs(p.v[1]), fabs(p.v[2])) + cs;
}

/* Primitive::sound_speed_squared, src/physics.rs:855-871 */
static double sound_speed_squared(const cdc_hydro_t *hy, const vec4 *pc, double phi) {
    if (hy->kind == CDC_ISOTHERMAL) return -phi / (hy->mach_number * hy->mach_number);
    return hy->gamma_law_index * pc->v[3] / pc->v[0];
}

/* intercell_flux: iso src/physics.rs:584-607, Euler src/physics.rs:776-803 */
static vec4 intercell_flux(const cdc_hydro_t *hy, const cdc_physics_t *ph,
                           const vec4 *lpc, const vec4 *lgx, const vec4 *lgy,
                           const vec4 *rpc, const vec4 *rgx, const vec4 *rgy,
                           double lnu, double rnu, double dx, double dy, double phi, int axis) {
    int nv = hy->kind == CDC_EULER ? 4 : 3;
    const vec4 *lg = axis == 0 ? lgx : lgy;
    const vec4 *rg = axis == 0 ? rgx : rgy;
    vec4 pl = v_add(*lpc, v_scale(*lg, 0.5, nv), nv);
    vec4 pr = v_sub(*rpc, v_scale(*rg, 0.5, nv), nv);
    double lam = ph->lambda_multiplier;
    double tau_x = 0.5 * (stress_field(lpc, lgx, lgy, lnu, lam * lnu, dx, dy, axis, 0) +
                          stress_field(rpc, rgx, rgy, rnu, lam * rnu, dx, dy, axis, 0));
    double tau_y = 0.5 * (stress_field(lpc, lgx, lgy, lnu, lam * lnu, dx, dy, axis, 1) +
                          stress_field(rpc, rgx, rgy, rnu, lam * rnu, dx, dy, axis, 1));
    vec4 f;
    if (hy->kind == CDC_ISOTHERMAL) {
        double cs2 = -phi / (hy->mach_number * hy->mach_number);
        f = iso_riemann_hlle(pl, pr, axis, cs2);
        vec4 visc = {{0.0, -tau_x, -tau_y, 0.0}};
        f = v_add(f, visc, 3);
    } else {
        double vx = 0.5 * (lpc->v[1] + rpc->v[1]);
        double vy = 0.5 * (lpc->v[2] + rpc->v[2]);
        vec4 visc = {{0.0, -tau_x, -tau_y, -(tau_x * vx + tau_y * vy)}};
        f = euler_riemann_hlle(pl, pr, axis, hy->gamma_law_index);
        f = v_add(f, visc, 4);
    }
    return f;
}

/* ------------------------------------------------------------------------- */
/* Oracle object                                                             */

static int g_threads = 0;

void oracle_set_threads(int n) { g_threads = n; }

int oracle_get_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

static void st_alloc(oracle_t *o, ostate *s) {
    s->u = (double *)calloc((size_t)o->nblk * o->ncell * o->nv, sizeof(double));
    s->ist = (double *)calloc((size_t)o->nblk * NITEM * o->nv, sizeof(double));
    s->oec = (double *)calloc((size_t)o->nblk * NITEM * 4, sizeof(double));
    s->tq = (double *)calloc((size_t)o->nblk * 2, sizeof(double));
}

static void st_free(ostate *s) {
    free(s->u);
    free(s->ist);
    free(s->oec);
    free(s->tq);
}

static void st_copy(const oracle_t *o, ostate *d, const ostate *s) {
    d->time = s->time;
    d->num = s->num;
    d->den = s->den;
    memcpy(d->u, s->u, sizeof(double) * (size_t)o->nblk * o->ncell * o->nv);
    memcpy(d->ist, s->ist, sizeof(double) * (size_t)o->nblk * NITEM * o->nv);
    memcpy(d->oec, s->oec, sizeof(double) * (size_t)o->nblk * NITEM * 4);
    memcpy(d->tq, s->tq, sizeof(double) * (size_t)o->nblk * 2);
}

oracle_t *oracle_create(const cdc_mesh_t *mesh, const cdc_hydro_t *hydro, const cdc_physics_t *phys,
                        const double *initial) {
    oracle_t *o = (oracle_t *)calloc(1, sizeof(oracle_t));
    o->mesh = *mesh;
    o->hydro = *hydro;
    o->phys = *phys;
    o->nv = hydro->kind == CDC_EULER ? 4 : 3;
    o->nb = mesh->num_blocks;
    o->bs = mesh->block_size;
    o->nblk = o->nb * o->nb;
    o->ncell = (size_t)o->bs * o->bs;
    size_t n = (size_t)o->nblk * o->ncell * o->nv;
    o->initial = (double *)malloc(n * sizeof(double));
    memcpy(o->initial, initial, n * sizeof(double));
    st_alloc(o, &o->st);
    for (int k = 0; k < 3; ++k) st_alloc(o, &o->work[k]);
    memcpy(o->st.u, initial, n * sizeof(double));
    o->st.time = 0.0;
    o->st.num = 0;
    o->st.den = 1;
    o->prim = (double *)malloc(n * sizeof(double));
    o->perr = (herr *)calloc((size_t)o->nblk, sizeof(herr));
    o->uerr = (herr *)calloc((size_t)o->nblk, sizeof(herr));
    return o;
}

void oracle_destroy(oracle_t *o) {
    if (!o) return;
    free(o->initial);
    st_free(&o->st);
    for (int k = 0; k < 3; ++k) st_free(&o->work[k]);
    free(o->prim);
    free(o->perr);
    free(o->uerr);
    free(o);
}

void oracle_set_physics(oracle_t *o, const cdc_physics_t *ph) { o->phys = *ph; }

void oracle_set_state(oracle_t *o, double time, int64_t num, int64_t den, const double *u,
                      const double *ist, const double *oec) {
    o->st.time = time;
    o->st.num = num;
    o->st.den = den;
    memcpy(o->st.u, u, sizeof(double) * (size_t)o->nblk * o->ncell * o->nv);
    if (ist) memcpy(o->st.ist, ist, sizeof(double) * (size_t)o->nblk * NITEM * o->nv);
    else memset(o->st.ist, 0, sizeof(double) * (size_t)o->nblk * NITEM * o->nv);
    if (oec) memcpy(o->st.oec, oec, sizeof(double) * (size_t)o->nblk * NITEM * 4);
    else memset(o->st.oec, 0, sizeof(double) * (size_t)o->nblk * NITEM * 4);
    memset(o->st.tq, 0, sizeof(double) * (size_t)o->nblk * 2);
}

void oracle_get_state(const oracle_t *o, double *time, int64_t *num, int64_t *den, double *u,
                      double *ist, double *oec) {
    if (time) *time = o->st.time;
    if (num) *num = o->st.num;
    if (den) *den = o->st.den;
    if (u) memcpy(u, o->st.u, sizeof(double) * (size_t)o->nblk * o->ncell * o->nv);
    if (ist) memcpy(ist, o->st.ist, sizeof(double) * (size_t)o->nblk * NITEM * o->nv);
    if (oec) memcpy(oec, o->st.oec, sizeof(double) * (size_t)o->nblk * NITEM * 4);
}

/* Per-block torque impulses on body 1 and 2 ([nblk][2]; extension). */
void oracle_get_torque(const oracle_t *o, double *tq) {
    memcpy(tq, o->st.tq, sizeof(double) * (size_t)o->nblk * 2);
}

/* State::max_signal_speed, src/state.rs:121-133 (max is order independent) */
static double state_max_signal_speed(const oracle_t *o, const ostate *s) {
    double a = 0.0;
    size_t n = (size_t)o->nblk * o->ncell;
    int nv = o->nv;
#pragma omp parallel for reduction(max : a) num_threads(oracle_get_threads()) schedule(static)
    for (size_t c = 0; c < n; ++c) {
        vec4 u = {{0, 0, 0, 0}};
        for (int k = 0; k < nv; ++k) u.v[k] = s->u[c * nv + k];
        double m = hydro_max_signal_speed(&o->hydro, u);
        a = fmax(m, a);
    }
    return a;
}

double oracle_max_signal_speed(const oracle_t *o) { return state_max_signal_speed(o, &o->st); }

/* Mesh::neighbor_block_indexes, src/mesh.rs:101-110 */
static void neighbor_blocks(int nb, int bi, int bj, int out[3][3]) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            int ni = (bi + nb + r - 1) % nb;
            int nj = (bj + nb + c - 1) % nb;
            out[r][c] = ni * nb + nj;
        }
}

/* try_block_primitive, src/scheme.rs:185-199: the first failing cell in C
 * order, with the cell centre as its position. */
static void block_primitive(oracle_t *o, const ostate *s, int b, double *xv, double *yv) {
    int bs = o->bs, nv = o->nv;
    int bi = b / o->nb, bj = b % o->nb;
    herr *e = &o->perr[b];
    e->code = 0;
    int have_geom = 0;
    for (int i = 0; i < bs; ++i)
        for (int j = 0; j < bs; ++j) {
            size_t c = ((size_t)b * o->ncell + (size_t)i * bs + j) * nv;
            vec4 u = {{0, 0, 0, 0}}, p = {{0, 0, 0, 0}};
            for (int k = 0; k < nv; ++k) u.v[k] = s->u[c + k];
            herr ce = {0, 0, 0, 0};
            if (try_to_primitive(&o->hydro, u, &p, &ce)) {
                if (!e->code) {
                    if (!have_geom) {
                        block_vertices(&o->mesh, bi, xv);
                        block_vertices(&o->mesh, bj, yv);
                        have_geom = 1;
                    }
                    *e = ce;
                    e->x = (xv[i] + xv[i + 1]) / 2.0;
                    e->y = (yv[j] + yv[j + 1]) / 2.0;
                }
                for (int k = 0; k < nv; ++k) o->prim[c + k] = 0.0;
                continue;
            }
            for (int k = 0; k < nv; ++k) o->prim[c + k] = p.v[k];
        }
}

typedef struct {
    vec4 *pe;  /* (bs+4)^2 extended primitives */
    vec4 *gx;  /* (bs+2) x (bs+4) */
    vec4 *gy;  /* (bs+4) x (bs+2) */
    vec4 *fx;  /* (bs+1) x bs */
    vec4 *fy;  /* bs x (bs+1) */
    double *xv, *yv, *xc, *yc;
} scratch_t;

/* compute_block_fluxes, src/scheme.rs:201-271, with the guard extension of
 * src/scheme.rs:126-127 (2 zones from the periodic 3x3 neighbours). */
static void block_fluxes(const oracle_t *o, int b, double time, const orbit_state *tbs, scratch_t *w) {
    const cdc_hydro_t *hy = &o->hydro;
    const cdc_physics_t *ph = &o->phys;
    int bs = o->bs, nv = o->nv, nb = o->nb;
    int bi = b / nb, bj = b % nb;
    int ne = bs + 4;
    int nbr[3][3];
    (void)time;
    neighbor_blocks(nb, bi, bj, nbr);

    /* ndarray_ops::extend_from_neighbor_arrays_2d(pn, 2, 2, 2, 2) */
    for (int ei = 0; ei < ne; ++ei) {
        int li = ei - 2;
        int r = li < 0 ? 0 : (li >= bs ? 2 : 1);
        int si = li - (r - 1) * bs;
        for (int ej = 0; ej < ne; ++ej) {
            int lj = ej - 2;
            int c = lj < 0 ? 0 : (lj >= bs ? 2 : 1);
            int sj = lj - (c - 1) * bs;
            const double *src = o->prim + ((size_t)nbr[r][c] * o->ncell + (size_t)si * bs + sj) * nv;
            vec4 p = {{0, 0, 0, 0}};
            for (int k = 0; k < nv; ++k) p.v[k] = src[k];
            w->pe[ei * ne + ej] = p;
        }
    }

    double pdt = ph->has_plm_density_trigger ? ph->plm_density_trigger : 0.0;
    /* gx = map_stencil3(pe, Axis(0), ...): (bs+2) x (bs+4) */
    for (int i = 0; i < bs + 2; ++i)
        for (int j = 0; j < ne; ++j) {
            const vec4 *a = &w->pe[i * ne + j], *bb = &w->pe[(i + 1) * ne + j], *c = &w->pe[(i + 2) * ne + j];
            double th = (a->v[0] < pdt || bb->v[0] < pdt || c->v[0] < pdt) ? 0.0 : ph->plm;
            w->gx[i * ne + j] = plm_gradient(th, a, bb, c, nv);
        }
    /* gy = map_stencil3(pe, Axis(1), ...): (bs+4) x (bs+2) */
    for (int i = 0; i < ne; ++i)
        for (int j = 0; j < bs + 2; ++j) {
            const vec4 *a = &w->pe[i * ne + j], *bb = &w->pe[i * ne + j + 1], *c = &w->pe[i * ne + j + 2];
            double th = (a->v[0] < pdt || bb->v[0] < pdt || c->v[0] < pdt) ? 0.0 : ph->plm;
            w->gy[i * (bs + 2) + j] = plm_gradient(th, a, bb, c, nv);
        }

    double dx = cell_spacing(&o->mesh), dy = cell_spacing(&o->mesh);
    double rs = softening_length(ph);
    block_vertices(&o->mesh, bi, w->xv);
    block_vertices(&o->mesh, bj, w->yv);
    for (int k = 0; k < bs; ++k) {
        w->xc[k] = (w->xv[k] + w->xv[k + 1]) / 2.0;
        w->yc[k] = (w->yv[k] + w->yv[k + 1]) / 2.0;
    }

/* cell_data (ci, cj), ci,cj in [0, bs+2): pc = pe[ci+1][cj+1], gx[ci][cj+1], gy[ci+1][cj] */
#define PC(ci, cj) (&w->pe[((ci) + 1) * ne + (cj) + 1])
#define GX(ci, cj) (&w->gx[(ci) * ne + (cj) + 1])
#define GY(ci, cj) (&w->gy[((ci) + 1) * (bs + 2) + (cj)])

    double x1 = tbs->p[0].x, y1 = tbs->p[0].y, x2 = tbs->p[1].x, y2 = tbs->p[1].y;
    double m1 = tbs->p[0].m, m2 = tbs->p[1].m;
    double gmi = hy->kind == CDC_EULER ? hy->gamma_law_index : 1.0;

    /* fx: faces (xv[i], yc[j]), L = cell (i-1, j), R = cell (i, j) */
    for (int i = 0; i <= bs; ++i)
        for (int j = 0; j < bs; ++j) {
            double fxp = w->xv[i], fyp = w->yc[j];
            const vec4 *lpc = PC(i, j + 1), *lgx = GX(i, j + 1), *lgy = GY(i, j + 1);
            const vec4 *rpc = PC(i + 1, j + 1), *rgx = GX(i + 1, j + 1), *rgy = GY(i + 1, j + 1);
            double phi = hy->kind == CDC_ISOTHERMAL ? gravitational_potential(tbs, fxp, fyp, rs) : 0.0;
            double lnu, rnu;
            if (ph->viscosity_model == CDC_VISC_CONSTANT_NU) {
                lnu = rnu = ph->viscosity;
            } else {
                double r1 = sqrt((fxp - x1) * (fxp - x1) + (fyp - y1) * (fyp - y1));
                double r2 = sqrt((fxp - x2) * (fxp - x2) + (fyp - y2) * (fyp - y2));
                double gpot = gravitational_potential(tbs, fxp, fyp, rs);
                double d1 = r1 + 1e-12 * dx, d2 = r2 + 1e-12 * dx;
                double twofreq = sqrt(m1 / (d1 * d1 * d1) + m2 / (d2 * d2 * d2));
                lnu = ph->viscosity * sound_speed_squared(hy, lpc, gpot) / twofreq / sqrt(gmi);
                rnu = ph->viscosity * sound_speed_squared(hy, rpc, gpot) / twofreq / sqrt(gmi);
            }
            w->fx[i * bs + j] = intercell_flux(hy, ph, lpc, lgx, lgy, rpc, rgx, rgy, lnu, rnu, dx, dy, phi, 0);
        }
    /* fy: faces (xc[i], yv[j]), L = cell (i, j-1), R = cell (i, j) */
    for (int i = 0; i < bs; ++i)
        for (int j = 0; j <= bs; ++j) {
            double fxp = w->xc[i], fyp = w->yv[j];
            const vec4 *lpc = PC(i + 1, j), *lgx = GX(i + 1, j), *lgy = GY(i + 1, j);
            const vec4 *rpc = PC(i + 1, j + 1), *rgx = GX(i + 1, j + 1), *rgy = GY(i + 1, j + 1);
            double phi = hy->kind == CDC_ISOTHERMAL ? gravitational_potential(tbs, fxp, fyp, rs) : 0.0;
            double lnu, rnu;
            if (ph->viscosity_model == CDC_VISC_CONSTANT_NU) {
                lnu = rnu = ph->viscosity;
            } else {
                double r1 = sqrt((fxp - x1) * (fxp - x1) + (fyp - y1) * (fyp - y1));
                double r2 = sqrt((fxp - x2) * (fxp - x2) + (fyp - y2) * (fyp - y2));
                double gpot = gravitational_potential(tbs, fxp, fyp, rs);
                double d1 = r1 + 1e-12 * dx, d2 = r2 + 1e-12 * dx;
                double twofreq = sqrt(m1 / (d1 * d1 * d1) + m2 / (d2 * d2 * d2));
                lnu = ph->viscosity * sound_speed_squared(hy, lpc, gpot) / twofreq / sqrt(gmi);
                rnu = ph->viscosity * sound_speed_squared(hy, rpc, gpot) / twofreq / sqrt(gmi);
            }
            w->fy[i * (bs + 1) + j] = intercell_flux(hy, ph, lpc, lgx, lgy, rpc, rgx, rgy, lnu, rnu, dx, dy, phi, 1);
        }
#undef PC
#undef GX
#undef GY
}

/* compute_block_updated_solution, src/scheme.rs:273-316 */
static int block_update(oracle_t *o, int b, const ostate *in, ostate *out, double time, double dt,
                        const orbit_state *tbs, scratch_t *w) {
    int bs = o->bs, nv = o->nv;
    double dx = cell_spacing(&o->mesh), dy = cell_spacing(&o->mesh);
    itemized ds;
    memset(&ds, 0, sizeof(ds));
    for (int i = 0; i < bs; ++i)
        for (int j = 0; j < bs; ++j) {
            size_t c = ((size_t)b * o->ncell + (size_t)i * bs + j) * nv;
            const vec4 *fxl = &w->fx[i * bs + j], *fxr = &w->fx[(i + 1) * bs + j];
            const vec4 *fyl = &w->fy[i * (bs + 1) + j], *fyr = &w->fy[i * (bs + 1) + j + 1];
            vec4 uc = {{0, 0, 0, 0}}, u0 = {{0, 0, 0, 0}}, du;
            for (int k = 0; k < nv; ++k) {
                du.v[k] = ((fxr->v[k] - fxl->v[k]) / dx + (fyr->v[k] - fyl->v[k]) / dy) * -dt;
                uc.v[k] = in->u[c + k];
                u0.v[k] = o->initial[c + k];
            }
            double x = w->xc[i], y = w->yc[j];
            itemized src = o->hydro.kind == CDC_EULER
                               ? euler_source_terms(&o->hydro, &o->phys, &o->mesh, uc, u0, x, y, time, dt, tbs)
                               : iso_source_terms(&o->phys, &o->mesh, uc, u0, x, y, dt, tbs);
            vec4 total = src.it[0];
            for (int it = 1; it < NITEM; ++it) total = v_add(total, src.it[it], nv);
            for (int it = 0; it < NITEM; ++it) ds.it[it] = v_add(ds.it[it], src.it[it], nv);
            for (int k = 0; k < nv; ++k) out->u[c + k] = uc.v[k] + du.v[k] + total.v[k];
        }
    for (int it = 0; it < NITEM; ++it) ds.it[it] = v_scale(ds.it[it], dx * dy, nv);
    double el[4];
    elements de[NITEM];
    oracle_orbital_elements(&o->phys, el);
    int rc = perturbation(&ds, tbs, el, de);
    if (rc) {
        o->uerr[b].code = rc;
        o->uerr[b].value = 0.0;
        o->uerr[b].x = 0.0;
        o->uerr[b].y = 0.0;
        return rc;
    }
    for (int it = 0; it < NITEM; ++it) {
        for (int k = 0; k < nv; ++k)
            out->ist[((size_t)b * NITEM + it) * nv + k] = in->ist[((size_t)b * NITEM + it) * nv + k] + ds.it[it].v[k];
        for (int k = 0; k < 4; ++k)
            out->oec[((size_t)b * NITEM + it) * 4 + k] = in->oec[((size_t)b * NITEM + it) * 4 + k] + de[it].e[k];
    }
    /* torque impulse on body k about the origin: x_k x (-dp), dp the gas
       momentum change from body k's gravity (no reference counterpart) */
    for (int k = 0; k < 2; ++k) {
        const vec4 *g = &ds.it[CDC_ITEM_GRAV1 + k];
        const double tau = tbs->p[k].y * g->v[1] - tbs->p[k].x * g->v[2];
        out->tq[(size_t)b * 2 + k] = in->tq[(size_t)b * 2 + k] + tau;
    }
    return 0;
}

static void fill_error(cdc_error_t *err, const herr *e, double time, const orbit_state *s) {
    if (!err) return;
    memset(err, 0, sizeof(*err));
    err->code = e->code;
    err->value = e->value;
    err->x = e->x;
    err->y = e->y;
    err->time = time;
    err->has_binary = 1;
    memcpy(err->binary, s, sizeof(*s));
    snprintf(err->message, sizeof(err->message), "%s",
             e->code == CDC_ERR_NEGATIVE_DENSITY    ? "negative surface density"
             : e->code == CDC_ERR_NEGATIVE_PRESSURE ? "negative gas pressure"
                                                    : "unbound orbital state");
}

/* try_advance_rk, src/scheme.rs:88-173 */
static int try_advance_rk(oracle_t *o, const ostate *in, ostate *out, double dt, cdc_error_t *err) {
    double time = in->time;
    orbit_state tbs;
    orbital_state_from_time(&o->phys, time, &tbs);
    int nthreads = oracle_get_threads();
    int bs = o->bs;

    /* P-phase */
#pragma omp parallel num_threads(nthreads)
    {
        double *xv = (double *)malloc(sizeof(double) * (bs + 1));
        double *yv = (double *)malloc(sizeof(double) * (bs + 1));
#pragma omp for schedule(dynamic, 1)
        for (int b = 0; b < o->nblk; ++b) block_primitive(o, in, b, xv, yv);
        free(xv);
        free(yv);
    }

    /* F- and U-phase.  Fluxes are block local (need_flux_communication()
     * is false, src/physics.rs:357-359), so each block's F and U run in one
     * task; a block fails if any of its 3x3 neighbours failed its P-phase
     * (try_join_3by3 order, src/scheme.rs:73-82, 126). */
#pragma omp parallel num_threads(nthreads)
    {
        scratch_t w;
        int ne = bs + 4;
        w.pe = (vec4 *)malloc(sizeof(vec4) * ne * ne);
        w.gx = (vec4 *)malloc(sizeof(vec4) * (bs + 2) * ne);
        w.gy = (vec4 *)malloc(sizeof(vec4) * ne * (bs + 2));
        w.fx = (vec4 *)malloc(sizeof(vec4) * (bs + 1) * bs);
        w.fy = (vec4 *)malloc(sizeof(vec4) * bs * (bs + 1));
        w.xv = (double *)malloc(sizeof(double) * (bs + 1));
        w.yv = (double *)malloc(sizeof(double) * (bs + 1));
        w.xc = (double *)malloc(sizeof(double) * bs);
        w.yc = (double *)malloc(sizeof(double) * bs);
#pragma omp for schedule(dynamic, 1)
        for (int b = 0; b < o->nblk; ++b) {
            int nbr[3][3];
            neighbor_blocks(o->nb, b / o->nb, b % o->nb, nbr);
            o->uerr[b].code = 0;
            for (int r = 0; r < 3 && !o->uerr[b].code; ++r)
                for (int c = 0; c < 3 && !o->uerr[b].code; ++c)
                    if (o->perr[nbr[r][c]].code) o->uerr[b] = o->perr[nbr[r][c]];
            if (o->uerr[b].code) continue;
            block_fluxes(o, b, time, &tbs, &w);
            block_update(o, b, in, out, time, dt, &tbs, &w);
        }
        free(w.pe); free(w.gx); free(w.gy); free(w.fx); free(w.fy);
        free(w.xv); free(w.yv); free(w.xc); free(w.yc);
    }

    /* join_all: the first Err in block order */
    for (int b = 0; b < o->nblk; ++b)
        if (o->uerr[b].code) {
            fill_error(err, &o->uerr[b], time, &tbs);
            return o->uerr[b].code;
        }
    out->time = in->time + dt;
    out->num = in->num + in->den;
    out->den = in->den;
    rat_norm(&out->num, &out->den);
    return 0;
}

/* WeightedAverage for State, src/state.rs:183-219: s1 <- s1 (1 - b) + s0 b */
static void weighted_average(oracle_t *o, ostate *s1, int64_t bn, int64_t bd, const ostate *s0) {
    double bf = (double)bn / (double)bd;
    double c1 = -bf + 1.0;
    size_t n = (size_t)o->nblk * o->ncell * o->nv;
#pragma omp parallel for num_threads(oracle_get_threads()) schedule(static)
    for (size_t k = 0; k < n; ++k) s1->u[k] = s1->u[k] * c1 + s0->u[k] * bf;
    size_t ni = (size_t)o->nblk * NITEM * o->nv;
    for (size_t k = 0; k < ni; ++k) s1->ist[k] = s1->ist[k] * c1 + s0->ist[k] * bf;
    size_t ne = (size_t)o->nblk * NITEM * 4;
    for (size_t k = 0; k < ne; ++k) s1->oec[k] = s1->oec[k] * c1 + s0->oec[k] * bf;
    for (size_t k = 0; k < (size_t)o->nblk * 2; ++k) s1->tq[k] = s1->tq[k] * c1 + s0->tq[k] * bf;
    s1->time = s1->time * c1 + s0->time * bf;
    int64_t on, od;
    rat_wavg(s1->num, s1->den, s0->num, s0->den, bn, bd, &on, &od);
    s1->num = on;
    s1->den = od;
}

/* One RK step of the given order from st into a work buffer; returns the
 * index of the work buffer holding the result, or -code on error. */
static int rk_step(oracle_t *o, const ostate *s0, int rk, double dt, cdc_error_t *err, int avoid) {
    int a = -1, b = -1;
    for (int k = 0; k < 3; ++k)
        if (k != avoid) { if (a < 0) a = k; else if (b < 0) b = k; }
    ostate *w1 = &o->work[a], *w2 = &o->work[b];
    int rc = try_advance_rk(o, s0, w1, dt, err);
    if (rc) return -rc;
    if (rk == 1) return a;
    rc = try_advance_rk(o, w1, w2, dt, err);
    if (rc) return -rc;
    if (rk == 2) {
        weighted_average(o, w2, 1, 2, s0);
        return b;
    }
    weighted_average(o, w2, 3, 4, s0);
    rc = try_advance_rk(o, w2, w1, dt, err);
    if (rc) return -rc;
    weighted_average(o, w1, 1, 3, s0);
    return a;
}

/* scheme::advance, src/scheme.rs:323-349 */
int oracle_advance(oracle_t *o, int fold, double *dt_out, cdc_error_t *err) {
    if (err) memset(err, 0, sizeof(*err));
    double a = state_max_signal_speed(o, &o->st);
    double dt = o->phys.cfl * cell_spacing(&o->mesh) / a;
    if (o->phys.has_dt_min) dt = fmax(dt, o->phys.dt_min);
    if (dt_out) *dt_out = dt;
    int rk = o->phys.rk_order;
    if (rk < 1 || rk > 3) return CDC_ERR_INVALID_ARGUMENT;
    const ostate *cur = &o->st;
    int cur_idx = -1;
    for (int f = 0; f < fold; ++f) {
        int r = rk_step(o, cur, rk, dt, err, cur_idx);
        if (r < 0) return -r;   /* o->st is untouched: transactional */
        cur_idx = r;
        cur = &o->work[r];
    }
    if (cur_idx >= 0) st_copy(o, &o->st, &o->work[cur_idx]);
    return 0;
}

int oracle_nv(const oracle_t *o) { return o->nv; }
