// cdc_device.h -- device arithmetic shared by the gfx950 kernels of the
// per-block Runge-Kutta update: refined FP64 reciprocal / square-root
// estimates, block geometry, PLM, primitive recovery, the max signal speed,
// the face flux (HLLE + viscous stress) and the cell source terms.
// Included by the HIP translation units only.
#pragma once

#include <hip/hip_runtime.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>

#include "cdc_internal.h"

namespace cdc {

// Element (v, c) of a device state of nv conserved variables over nn cells
// (c = gi * N + gj): variables 0 and 1 interleaved in a pair plane [nn][2],
// then variables 2 and 3 in a second pair plane (gamma-law gas) or variable
// 2 in a plane [nn] of its own (isothermal gas), so that a lane loads a
// cell's pair of values in one 16-byte access.
__host__ __device__ __forceinline__ int64_t sidx(int nv, int v, int64_t nn, int64_t c) {
    return v < 2 ? 2 * c + v : 2 * nn + (nv == 4 ? 2 * c + (v - 2) : c);
}

// ---------------------------------------------------------------------------
// Compact list of the itemized-source components that can be nonzero; the
// others (grav mass, iso cooling, Euler fake_mass) are identically zero.

template <int NV> struct SumMap;
template <> struct SumMap<3> {
    static constexpr int n = 16;
    __host__ __device__ static constexpr int item(int s) {
        return s < 3 ? CDC_ITEM_SINK1 : s < 6 ? CDC_ITEM_SINK2 : s < 8 ? CDC_ITEM_GRAV1
             : s < 10 ? CDC_ITEM_GRAV2 : s < 13 ? CDC_ITEM_BUFFER : CDC_ITEM_FAKE_MASS;
    }
    __host__ __device__ static constexpr int comp(int s) {
        return s < 3 ? s : s < 6 ? s - 3 : s < 8 ? s - 5 : s < 10 ? s - 7 : s < 13 ? s - 10 : s - 13;
    }
};
template <> struct SumMap<4> {
    static constexpr int n = 22;
    __host__ __device__ static constexpr int item(int s) {
        return s < 4 ? CDC_ITEM_SINK1 : s < 8 ? CDC_ITEM_SINK2 : s < 11 ? CDC_ITEM_GRAV1
             : s < 14 ? CDC_ITEM_GRAV2 : s < 18 ? CDC_ITEM_BUFFER : CDC_ITEM_COOLING;
    }
    __host__ __device__ static constexpr int comp(int s) {
        return s < 4 ? s : s < 8 ? s - 4 : s < 11 ? s - 7 : s < 14 ? s - 10 : s < 18 ? s - 14 : s - 18;
    }
};

namespace dev {

// Device memory reached through a pointer loaded from the per-run argument
// array is not known to the compiler to be global memory; these casts let it
// use global (not flat) loads and stores, which wait on vmcnt alone instead
// of also on the LDS / scalar counter.
typedef __attribute__((address_space(1))) const double gcdouble;
typedef __attribute__((address_space(1))) double gdouble;
__device__ __forceinline__ gcdouble *gptr(const double *p) { return (gcdouble *)p; }
__device__ __forceinline__ gdouble *gptr(double *p) { return (gdouble *)p; }

// ---- refined hardware estimates -------------------------------------------

// 1 / x: one Newton step from the estimate
__device__ __forceinline__ double frcp(double x) {
    const double y = __builtin_amdgcn_rcp(x);
    const double e = __builtin_fma(-x, y, 1.0);
    return __builtin_fma(y, e, y);
}

// 1 / sqrt(x), x > 0: one step of the coupled Newton iteration on
// (g ~ sqrt x, h ~ 1/(2 sqrt x)); the step's result 2 (h + h r) is formed
// as y + y r (y = 2 h exactly: the same bits, one multiplication fewer)
__device__ __forceinline__ double frsq(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * y, g = x * y;
    const double r = __builtin_fma(-h, g, 0.5);
    return __builtin_fma(y, r, y);
}

// sqrt(x), x >= 0: one coupled step, then one residual correction of g
__device__ __forceinline__ double fsqrt(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double h = 0.5 * y, g = x * y;
    const double r = __builtin_fma(-h, g, 0.5);
    h = __builtin_fma(h, r, h);
    g = __builtin_fma(g, r, g);
    const double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return x > 0.0 ? g : x;
}

// sqrt(x + 1e-300), x >= 0, without the zero test: the tiny addend changes
// no x above 2e-284 (it is below half an ulp there) and keeps x = 0 finite
// (sqrt(1e-300) = 1e-150), so a squared distance needs no select.  Formed
// as x times its refined 1/sqrt: a few ulp from the rounded root, where
// fsqrt's residual correction costs two operations more (the alpha face
// factor, whose distances enter the viscosity only)
__device__ __forceinline__ double fsqrt_nn(double x) {
    x += 1e-300;
    return x * frsq(x);
}

// Block geometry exactly as Mesh::block_vertices / cell_centers
// (src/mesh.rs:45-84): linspace(start, start + L, bs + 1), centres as
// adjacent means.  No contraction, so the coordinates match the CPU's bits.
struct Axis {
    double a, step;
    __device__ __forceinline__ double v(int k) const {
#pragma clang fp contract(off)
        return a + step * (double)k;
    }
    __device__ __forceinline__ double c(int k) const {
#pragma clang fp contract(off)
        return (v(k) + v(k + 1)) / 2.0;
    }
};

__device__ __forceinline__ Axis block_axis(double R, double L, int bi, int bs) {
#pragma clang fp contract(off)
    Axis ax;
    ax.a = -R + (double)bi * L;
    const double b = ax.a + L;
    ax.step = (b - ax.a) / (double)bs;
    return ax;
}

// minmod of three values: the median of (0, min, max), i.e. the smallest
// magnitude when all three share a sign and 0 otherwise
__device__ __forceinline__ double minmod3(double a, double b, double c) {
    const double mn = fmin(a, fmin(b, c));
    const double mx = fmax(a, fmax(b, c));
    return fmax(mn, fmin(mx, 0.0));
}

// godunov-core plm_gradient: minmod(theta (b - a), (c - a) / 2, theta (c - b))
__device__ __forceinline__ double plm(double th, double a, double b, double c) {
    return minmod3(th * (b - a), 0.5 * (c - a), th * (c - b));
}

// try_to_primitive: iso src/physics.rs:489-503, Euler src/physics.rs:631-660.
// Also returns 1/Sigma of the recovered state in `id`, the pressure before
// the pressure floor in `praw` (Euler; the source terms' recovery of the
// same state has no floors), and in `val` the offending value when the
// result is an error code.  The floors compare against
// StageConsts::*_floor_cmp (-inf when unset) and the error is chosen by
// selects, so the recovery has no per-cell branch but the rare floor.
// DFLOOR = false: the run has no density floor (the test is left out).
template <int NV, bool DFLOOR = true>
__device__ __forceinline__ int try_to_primitive(const StageConsts &k, const double *u, double *p,
                                                double &val, double &id, double &praw) {
    if constexpr (NV == 3) {
        double d = u[0], sx = u[1], sy = u[2];
        if (DFLOOR && d < k.density_floor_cmp) {
            sx = 0.0; sy = 0.0; d = k.density_floor;
        }
        id = frcp(d);
        p[0] = d; p[1] = sx * id; p[2] = sy * id;
        praw = 0.0;
        val = d;
        return d < 0.0 ? CDC_ERR_NEGATIVE_DENSITY : 0;
    } else {
        const double d = u[0];
        id = frcp(d);
        const double vx = u[1] * id, vy = u[2] * id;
        p[0] = d; p[1] = vx; p[2] = vy;
        p[3] = (u[3] - 0.5 * d * (vx * vx + vy * vy)) * k.gm1;
        praw = p[3];
        if (DFLOOR && d < k.density_floor_cmp) {
            p[0] = k.density_floor; p[1] = 0.0; p[2] = 0.0; p[3] = k.density_floor_pressure;
            id = frcp(k.density_floor);
        }
        // negative density first; the pressure floor then the pressure check
        const bool nd = p[0] < 0.0;
        const double pf = p[3] < k.pressure_floor_cmp ? k.pressure_floor_cmp : p[3];
        p[3] = pf;
        val = nd ? p[0] : pf;
        return nd ? CDC_ERR_NEGATIVE_DENSITY : (pf <= 0.0 ? CDC_ERR_NEGATIVE_PRESSURE : 0);
    }
}

template <int NV>
__device__ __forceinline__ int try_to_primitive(const StageConsts &k, const double *u, double *p,
                                                double &val, double &id) {
    double praw;
    return try_to_primitive<NV, true>(k, u, p, val, id, praw);
}

// Hydrodynamics::max_signal_speed of a conserved state: iso
// max(|vx|, |vy|) (src/physics.rs:521-526), Euler max(|vx|, |vy|) + c_s.
// Isothermal: max(|Sx|, |Sy|) / |Sigma|, one IEEE division, equal to the max
// of the two correctly rounded quotients (rounding is monotonic).  Euler:
// the refined reciprocal and square root (within 11 and 0.5 ulp).
template <int NV>
__device__ __forceinline__ double max_signal_speed(const double *u, double gm) {
    if constexpr (NV == 3) {
        return fmax(fabs(u[1]), fabs(u[2])) / fabs(u[0]);
    } else {
        const double d = u[0];
        const double id = frcp(d);
        const double vx = u[1] * id, vy = u[2] * id;
        const double p = (u[3] - 0.5 * d * (vx * vx + vy * vy)) * (gm - 1.0);
        return fmax(fabs(vx), fabs(vy)) + fsqrt(gm * p * id);
    }
}

// One face: reconstruction, viscous stress, HLLE.  axis 0 = x, 1 = y.
// eL/eR index the primitive tile (and the cell sound-speed plane C), rL/rR
// the gradient ring.
template <int NV, int EXY, int RXY>
__device__ __forceinline__ void face_flux(const StageArgs &A, int axis, const double *P, const double *C,
                                          const double *GX, const double *GY, int eL, int eR, int rL,
                                          int rR, double fxp, double fyp, double *F) {
    const StageConsts &k = A.k;
    const double *GN = axis ? GY : GX;
    double pL[NV], pR[NV], pl[NV], pr[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        pL[v] = P[v * EXY + eL];
        pR[v] = P[v * EXY + eR];
        pl[v] = pL[v] + GN[v * RXY + rL] * 0.5;
        pr[v] = pR[v] - GN[v * RXY + rR] * 0.5;
    }

    // Plummer potential at the face: isothermal c_s^2 = -phi / M^2
    const Orbit &o = A.orb;
    const double d1x = fxp - o.x1, d1y = fyp - o.y1;
    const double d2x = fxp - o.x2, d2y = fyp - o.y2;
    const double r1s = d1x * d1x + d1y * d1y;
    const double r2s = d2x * d2x + d2y * d2y;
    double cs2f = 0.0;
    if constexpr (NV == 3) cs2f = (o.m1 * frsq(r1s + k.rs2) + o.m2 * frsq(r2s + k.rs2)) * k.inv_mach2;

    // viscous stress tau = (sigma_L + sigma_R) / 2 (src/physics.rs:322-340)
    double tau_x = 0.0, tau_y = 0.0;
    if (k.viscosity_model != CDC_VISC_CONSTANT_NU || k.visc != 0.0) {
        double nuL, nuR;
        if (k.viscosity_model == CDC_VISC_CONSTANT_NU) {
            nuL = nuR = k.visc;
        } else {
            // alpha c_s^2 / sqrt(gamma) / sqrt(m1/q1^3 + m2/q2^3), q = r + 1e-12 dx
            const double q1 = fsqrt(r1s) + k.eps_dx;
            const double q2 = fsqrt(r2s) + k.eps_dx;
            const double c1 = q1 * q1 * q1, c2 = q2 * q2 * q2;
            const double itwofreq = frsq((o.m1 * c2 + o.m2 * c1) * frcp(c1 * c2));
            if constexpr (NV == 3) {
                nuL = nuR = k.visc * cs2f * itwofreq;
            } else {
                nuL = k.visc * C[eL] * itwofreq;
                nuR = k.visc * C[eR] * itwofreq;
            }
        }
        const double gxvxL = GX[1 * RXY + rL] * k.idx, gxvyL = GX[2 * RXY + rL] * k.idx;
        const double gyvxL = GY[1 * RXY + rL] * k.idy, gyvyL = GY[2 * RXY + rL] * k.idy;
        const double gxvxR = GX[1 * RXY + rR] * k.idx, gxvyR = GX[2 * RXY + rR] * k.idx;
        const double gyvxR = GY[1 * RXY + rR] * k.idy, gyvyR = GY[2 * RXY + rR] * k.idy;
        const double offL = gxvyL + gyvxL, offR = gxvyR + gyvxR;
        const double divL = gxvxL + gyvyL, divR = gxvxR + gyvyR;
        const double nsL = pL[0] * nuL, nsR = pR[0] * nuR;
        // normal-normal component (sxx on an x-face, syy on a y-face) and the
        // off-diagonal one, by selects so that both face kinds share one path
        const double gnnL = axis ? gyvyL : gxvxL, gttL = axis ? gxvxL : gyvyL;
        const double gnnR = axis ? gyvyR : gxvxR, gttR = axis ? gxvxR : gyvyR;
        const double snnL = nsL * ((4.0 / 3.0) * gnnL - (2.0 / 3.0) * gttL + k.lam * divL);
        const double snnR = nsR * ((4.0 / 3.0) * gnnR - (2.0 / 3.0) * gttR + k.lam * divR);
        const double tau_n = 0.5 * (snnL + snnR);
        const double tau_o = 0.5 * (nsL * offL + nsR * offR);
        tau_x = axis ? tau_o : tau_n;
        tau_y = axis ? tau_n : tau_o;
    }

    // HLLE (Toro ch. 10), outer wave speeds v_n -/+ c_s
    const double vnl = axis ? pl[2] : pl[1], vnr = axis ? pr[2] : pr[1];
    double ul[NV], ur[NV], fl[NV], fr[NV], csl, csr, ppl, ppr;
    ul[0] = pl[0]; ul[1] = pl[0] * pl[1]; ul[2] = pl[0] * pl[2];
    ur[0] = pr[0]; ur[1] = pr[0] * pr[1]; ur[2] = pr[0] * pr[2];
    if constexpr (NV == 3) {
        csl = csr = fsqrt(cs2f);
        ppl = pl[0] * cs2f;
        ppr = pr[0] * cs2f;
    } else {
        ul[3] = 0.5 * pl[0] * (pl[1] * pl[1] + pl[2] * pl[2]) + pl[3] * k.inv_gm1;
        ur[3] = 0.5 * pr[0] * (pr[1] * pr[1] + pr[2] * pr[2]) + pr[3] * k.inv_gm1;
        const double gpl = k.gm * pl[3], gpr = k.gm * pr[3];
        csl = gpl * frsq(gpl * pl[0]);      // sqrt(gamma p / Sigma)
        csr = gpr * frsq(gpr * pr[0]);
        ppl = pl[3];
        ppr = pr[3];
        fl[3] = (ul[3] + ppl) * vnl;
        fr[3] = (ur[3] + ppr) * vnr;
    }
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        fl[v] = ul[v] * vnl;
        fr[v] = ur[v] * vnr;
    }
    fl[1] = axis ? fl[1] : fl[1] + ppl;
    fl[2] = axis ? fl[2] + ppl : fl[2];
    fr[1] = axis ? fr[1] : fr[1] + ppr;
    fr[2] = axis ? fr[2] + ppr : fr[2];
    const double ap = fmax(fmax(vnl + csl, vnr + csr), 0.0);
    const double am = fmin(fmin(vnl - csl, vnr - csr), 0.0);
    const double inv = frcp(ap - am);
    const double apam = ap * am;
#pragma unroll
    for (int v = 0; v < NV; ++v) F[v] = (fl[v] * ap - fr[v] * am - (ul[v] - ur[v]) * apam) * inv;
    F[1] -= tau_x;
    F[2] -= tau_y;
    if constexpr (NV == 4) {
        const double vx = 0.5 * (pL[1] + pR[1]);
        const double vy = 0.5 * (pL[2] + pR[2]);
        F[3] -= tau_x * vx + tau_y * vy;
    }
}

// Where a cell sits relative to the two bodies, and its sink kernel rates.
struct SinkGeom {
    double ex1, ey1, ex2, ey2, q1, q2, sr1, sr2;
};

// The sink geometry of a cell at (x, y): its offsets from the bodies, their
// squares' sums and the two sink kernel rates (1 - r^2/s^2)^2 inside the sink
// radius.
__device__ __forceinline__ SinkGeom sink_geom(const StageConsts &k, const Orbit &o, double x, double y) {
    SinkGeom g;
    g.ex1 = x - o.x1; g.ey1 = y - o.y1; g.ex2 = x - o.x2; g.ey2 = y - o.y2;
    g.q1 = g.ex1 * g.ex1 + g.ey1 * g.ey1;
    g.q2 = g.ex2 * g.ex2 + g.ey2 * g.ey2;
    g.sr1 = 0.0; g.sr2 = 0.0;
    if (g.q1 < k.sink_s2) { const double w = 1.0 - g.q1 * k.inv_sink_s2; g.sr1 = k.sink_rate * (w * w); }
    if (g.q2 < k.sink_s2) { const double w = 1.0 - g.q2 * k.inv_sink_s2; g.sr2 = k.sink_rate * (w * w); }
    return g;
}

// The two sink items (src/physics.rs:548-569, 717-744): impartial, or
// torque-free / acceleration-free with the gas's velocity relative to the
// body along the line of centres.  it[SINK1], it[SINK2] (times dt).
template <int NV>
__device__ __forceinline__ void sink_items(const StageConsts &k, const Orbit &o, double dt, const double *u,
                                           double vx, double vy, double id, double p0, const SinkGeom &g,
                                           double it[7][NV]) {
    const double sd = u[0], sr1 = g.sr1, sr2 = g.sr2;
    double s1[NV], s2[NV];
    if (k.sink_model == CDC_SINK_IMPARTIAL) {
#pragma unroll
        for (int v = 0; v < NV; ++v) { s1[v] = u[v] * -sr1; s2[v] = u[v] * -sr2; }
    } else {
        const double vx1 = k.one_body ? 0.0 : o.vx1, vy1 = k.one_body ? 0.0 : o.vy1;
        const double vx2 = k.one_body ? 0.0 : o.vx2, vy2 = k.one_body ? 0.0 : o.vy2;
        const double ir1 = frsq(g.q1), ir2 = frsq(g.q2);
        const double r1x = g.ex1 * ir1, r1y = g.ey1 * ir1, r2x = g.ex2 * ir2, r2y = g.ey2 * ir2;
        const double dv1 = (vx - vx1) * r1x + (vy - vy1) * r1y;
        const double dv2 = (vx - vx2) * r2x + (vy - vy2) * r2y;
        const double vx1c = dv1 * r1x + vx1, vy1c = dv1 * r1y + vy1;
        const double vx2c = dv2 * r2x + vx2, vy2c = dv2 * r2y + vy2;
        s1[0] = sd * -sr1; s1[1] = (sd * vx1c) * -sr1; s1[2] = (sd * vy1c) * -sr1;
        s2[0] = sd * -sr2; s2[1] = (sd * vx2c) * -sr2; s2[2] = (sd * vy2c) * -sr2;
        if constexpr (NV == 4) {
            const double e0 = p0 * (id * k.inv_gm1);
            s1[3] = (sd * e0 + 0.5 * sd * (vx1c * vx1c + vy1c * vy1c)) * -sr1;
            s2[3] = (sd * e0 + 0.5 * sd * (vx2c * vx2c + vy2c * vy2c)) * -sr2;
        }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        it[CDC_ITEM_SINK1][v] = s1[v] * dt;
        it[CDC_ITEM_SINK2][v] = s2[v] * dt;
    }
}

// Isothermal / Euler source_terms (src/physics.rs:528-582, 683-774) with
// Physics::source_terms (src/physics.rs:423-457).  `sd, vx, vy, id, p0` is
// the primitive recovery of u without floors; `brate` the static buffer
// rate; `cool` the beta-cooling frequency Omega, or, when A.cool_factor, the
// relaxation factor exp(-dt Omega / beta) itself.  it[item][v].
// `cool`: the cooling law when it is fixed at compile time (StageCooling);
// kCoolAny selects it at run time from the stage arguments.
// `o` is the stage's orbit (A.orb, or a copy the caller holds in registers).
// With LAZY_SINKS the sink items are not written: `g` returns what
// sink_items needs, and a caller evaluates them only where a rate is
// nonzero (elsewhere they are +-0).
template <int NV, int COOL = kCoolAny, bool LAZY_SINKS = false>
__device__ __forceinline__ void cell_sources(const StageArgs &A, const Orbit &o, const double *u, const double *u0,
                                             double x, double y, double vx, double vy, double id, double p0,
                                             double brate, double cool, double it[7][NV], SinkGeom &g) {
    const StageConsts &k = A.k;
    const double dt = A.dt;
    const double sd = u[0];

    // Plummer gravity of each body: a = -m d / (r^2 + r_s^2)^(3/2)
    g = sink_geom(k, o, x, y);
    const double ex1 = g.ex1, ey1 = g.ey1, ex2 = g.ex2, ey2 = g.ey2, q1 = g.q1, q2 = g.q2;
    // (the items are impulses: the factor -m dt Sigma / (r^2 + r_s^2)^(3/2)
    // is formed once, dt folded in before the offsets; a rounding-level
    // regrouping of the reference's f dt)
    const double y1 = frsq(q1 + k.rs2), y2 = frsq(q2 + k.rs2);
    const double g1 = (o.m1 * -dt) * (y1 * y1 * y1) * sd, g2 = (o.m2 * -dt) * (y2 * y2 * y2) * sd;
    const double fx1 = g1 * ex1, fy1 = g1 * ey1, fx2 = g2 * ex2, fy2 = g2 * ey2;

    if constexpr (!LAZY_SINKS) sink_items<NV>(k, o, dt, u, vx, vy, id, p0, g, it);
    const double bfac = -dt * brate;
#pragma unroll
    for (int v = 0; v < NV; ++v) it[CDC_ITEM_BUFFER][v] = (u[v] - u0[v]) * bfac;
    it[CDC_ITEM_GRAV1][0] = 0.0; it[CDC_ITEM_GRAV1][1] = fx1; it[CDC_ITEM_GRAV1][2] = fy1;
    it[CDC_ITEM_GRAV2][0] = 0.0; it[CDC_ITEM_GRAV2][1] = fx2; it[CDC_ITEM_GRAV2][2] = fy2;

    if constexpr (NV == 3) {
#pragma unroll
        for (int v = 0; v < 3; ++v) it[CDC_ITEM_COOLING][v] = 0.0;
        if (u[0] < u0[0] * k.fake_threshold) {
#pragma unroll
            for (int v = 0; v < 3; ++v) it[CDC_ITEM_FAKE_MASS][v] = 0.0;
        } else {
#pragma unroll
            for (int v = 0; v < 3; ++v) it[CDC_ITEM_FAKE_MASS][v] = u0[v] * k.fake_rate * dt;
        }
    } else {
        it[CDC_ITEM_GRAV1][3] = fx1 * vx + fy1 * vy;
        it[CDC_ITEM_GRAV2][3] = fx2 * vx + fy2 * vy;
        const double e0 = p0 * (id * k.inv_gm1);
        double ec;
        const bool beta = COOL == kCoolAny ? k.cooling_model == CDC_COOLING_BETA : COOL != kCoolT4;
        if (beta) {
            if constexpr (COOL != kCoolT4) {
                const double ib0 = frcp(u0[0]);
                const double bvx = u0[1] * ib0, bvy = u0[2] * ib0;
                const double bp = (u0[3] - 0.5 * u0[0] * (bvx * bvx + bvy * bvy)) * k.gm1;
                const double et = bp * (ib0 * k.inv_gm1);
                double f;
                if constexpr (COOL == kCoolBetaFactor) f = cool;
                else if constexpr (COOL == kCoolBetaExp) f = exp(-dt * cool * k.cooling_inv_beta);
                else f = A.cool_factor ? cool : exp(-dt * cool * k.cooling_inv_beta);
                ec = et + (e0 - et) * f;
            }
        } else {
            if constexpr (COOL == kCoolAny || COOL == kCoolT4) {
                const double sdi = k.density_index == 1 ? sd : (k.density_index == 2 ? sd * sd : pow(sd, (double)k.density_index));
                const double a = k.cooling_coefficient / sdi * k.cooling_f;
                ec = e0 * pow(1.0 + 3.0 * a * (e0 * e0 * e0) * dt, -1.0 / 3.0);
            }
        }
        if (k.has_mach_ceiling) ec = fmax(ec, 0.5 * (vx * vx + vy * vy) * k.mach_ceiling_coef);
        // uc = to_conserved(Sigma, vx, vy, Sigma ec (gamma - 1)); cooling = uc - u
        const double pc = sd * ec * k.gm1;
        it[CDC_ITEM_COOLING][0] = sd - u[0];
        it[CDC_ITEM_COOLING][1] = sd * vx - u[1];
        it[CDC_ITEM_COOLING][2] = sd * vy - u[2];
        it[CDC_ITEM_COOLING][3] = (0.5 * sd * (vx * vx + vy * vy) + pc * k.inv_gm1) - u[3];
#pragma unroll
        for (int v = 0; v < 4; ++v) it[CDC_ITEM_FAKE_MASS][v] = 0.0;
    }
}

}  // namespace dev

// Entry e of the strip-edge columns of state u of nv variables (StageArgs::ec
// layout [nv][N][N / 64][4]: per variable and row, the columns
// 64 k - 2 .. 64 k + 1 of boundary k, periodic).
__device__ __forceinline__ double edge_entry(const double *u, int64_t e, int N, int nv) {
    const int NK = N / 64;
    const int c = (int)(e & 3);
    const int64_t rest = e >> 2;
    const int k = (int)(rest % NK);
    int col = 64 * k - 2 + c;
    col = col < 0 ? col + N : (col >= N ? col - N : col);
    const int64_t vr = rest / NK;   // v * N + row
    const int v = (int)(vr / N);
    return u[sidx(nv, v, (int64_t)N * N, (vr - (int64_t)v * N) * N + col)];
}

// ---------------------------------------------------------------------------
// Launch geometry and the per-block finalize shared by the stage kernels

// Tile index of workgroup wg: consecutive tiles go to one XCD (workgroups
// are dealt round-robin over the 8 XCDs), so halo re-reads of neighbouring
// tiles hit the same L2.
__device__ __forceinline__ int xcd_tile(int wg, int ntiles) {
    return (ntiles % 8 == 0) ? (wg % 8) * (ntiles / 8) + wg / 8 : wg;
}

// Run index of global tile (or block) t: the last r with begin[r] <= t.
__device__ __forceinline__ int run_of(const int32_t *__restrict__ begin, int nruns, int t) {
    int lo = 0, hi = nruns - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begin[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// A stage that already failed (err_flag holds an earlier stage id) skips the
// remaining stages of the call.
__device__ __forceinline__ bool stage_skipped(const StageArgs &A) {
    const int ef = *(volatile const int32_t *)A.err_flag;
    return ef != 0 && ef < A.stage_id;
}

// The per-block tail of a stage, one wave per block: the block's tile sums
// in tile order, scaled by dx*dy, the four orbital-element perturbations
// (ItemizedChange::perturbation, src/state.rs:258-282) one per lane, and the
// RK combine of integrated_source_terms / orbital_elements_change
// (src/scheme.rs:307-315, src/state.rs:195-197).
// A wave that only waits on memory: every load that does not depend on the
// tile sums (the error flag, the block's current and base ist / oec / torque
// values) is issued with the tile-sum loads, so the block costs two memory
// round trips (loads, then stores) around the perturbation arithmetic.
// One wave per block, lane = lane index; every wave of the workgroup calls
// this (an inactive one, past the last block, only takes the two barriers).
// Scratch in LDS: s_ds[7 nv], s_de[7 * 4], *s_perr, per wave.
constexpr int kFinLds = 7 * 4 + 7 * 4 + 1;   // doubles of per-wave scratch (nv <= 4)

template <int NV>
__device__ __forceinline__ void finalize_block(const StageArgs &A, const int b, const int tpb, const int lane,
                                               bool active, double *s_ds, double *s_de, int *s_perr) {
    constexpr int NS = SumMap<NV>::n, NR = NS + 1;
    static_assert(NS + 1 + 7 * NV <= 64, "one wave per block");
    const int ef = active ? *(volatile const int32_t *)A.err_flag : 0;
    const size_t oi = (size_t)b * 7 * NV, oe = (size_t)b * 7 * 4;
    double ist_c = 0.0, ist_b = 0.0, oec_c = 0.0, oec_b = 0.0, tq_c = 0.0, tq_b = 0.0;
    double bmax = 0.0;   // lane NS: the block's max signal speed (when A.amax_out)
    if (active) {
        if (lane < 7 * 4) {
            if (lane < 7 * NV) {
                ist_c = A.ist_c[oi + lane];
                if (A.read_base) ist_b = A.ist_b[oi + lane];
            }
            oec_c = A.oec_c[oe + lane];
            if (A.read_base) oec_b = A.oec_b[oe + lane];
        } else if (lane >= 32 && lane < 34) {
            tq_c = A.tq_c[(size_t)b * 2 + (lane - 32)];
            if (A.read_base) tq_b = A.tq_b[(size_t)b * 2 + (lane - 32)];
        }
        if (lane < NS || (lane == NS && A.amax_out)) {
            double t = 0.0;
            // tiles in order; unrolled so that the loads are in flight together
            const auto *ts = dev::gptr(A.tile_sums) + (size_t)b * tpb * NR + lane;
            int j = 0;
            const bool mx = lane == NS;   // the max-speed column: maxima, not sums
            for (; j + 8 <= tpb; j += 8) {
                double v[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) v[q] = ts[(size_t)(j + q) * NR];
#pragma unroll
                for (int q = 0; q < 8; ++q) t = mx ? fmax(t, v[q]) : t + v[q];
            }
            for (; j < tpb; ++j) t = mx ? fmax(t, ts[(size_t)j * NR]) : t + ts[(size_t)j * NR];
            if (!mx) s_ds[SumMap<NV>::item(lane) * NV + SumMap<NV>::comp(lane)] = t * A.area;
            else bmax = t;
        } else if (lane > NS && lane < NS + 1 + 7 * NV) {
            // structurally zero components
            const int z = lane - NS - 1;
            bool nonzero = false;
            for (int s = 0; s < NS; ++s) nonzero |= (SumMap<NV>::item(s) * NV + SumMap<NV>::comp(s)) == z;
            if (!nonzero) s_ds[z] = 0.0;
        }
    }
    if (lane == 0) *s_perr = 0;
    // a run whose earlier stage already failed (err_flag holds an earlier
    // stage id) writes nothing; the wave stays for the barriers (the waves
    // of one workgroup may belong to different runs)
    if (ef != 0 && ef < A.stage_id) active = false;
    __syncthreads();
    if (active) {
        if (lane < 4) {
            const int item = lane == 0 ? CDC_ITEM_SINK1 : lane == 1 ? CDC_ITEM_SINK2 : lane == 2 ? CDC_ITEM_GRAV1 : CDC_ITEM_GRAV2;
            const double *d = s_ds + item * NV;
            double p[4];
            const bool ok = (lane % 2 == 0) ? perturb(A.orb, -d[0], 0.0, -d[1], 0.0, -d[2], 0.0, p)
                                            : perturb(A.orb, 0.0, -d[0], 0.0, -d[1], 0.0, -d[2], p);
            for (int c = 0; c < 4; ++c) s_de[item * 4 + c] = ok ? p[c] - A.el[c] : 0.0;
            if (!ok) atomicOr(s_perr, 1);
        } else if (lane < 4 + 3 * 4) {
            const int z = lane - 4;   // buffer, cooling, fake_mass items: no orbital change
            s_de[(CDC_ITEM_BUFFER + z / 4) * 4 + (z % 4)] = 0.0;
        }
    }
    __syncthreads();
    if (!active) return;
    if (lane == NS && bmax > 0.0) {
        // after the last barrier, so that no wait covers this atomic.  Speeds
        // are >= 0, whose bit patterns order like their values: an integer
        // maximum, exact and independent of the blocks' order
        atomicMax((unsigned long long *)A.amax_out, (unsigned long long)__double_as_longlong(bmax));
    }
    if (*s_perr) {
        if (lane == 0) {
            A.perr[b] = CDC_ERR_UNBOUND_ORBIT;
            atomicCAS(A.err_flag, 0, A.stage_id);
        }
        return;
    }
    if (lane == 0) A.perr[b] = 0;
    if (lane >= 32 && lane < 34) {
        // torque impulse on body k about the origin: x_k x (-dp_k), dp_k the
        // gas momentum change from body k's gravity (no reference counterpart)
        const int k = lane - 32;
        const double *g = s_ds + (CDC_ITEM_GRAV1 + k) * NV;
        const double x = k == 0 ? A.orb.x1 : A.orb.x2, y = k == 0 ? A.orb.y1 : A.orb.y2;
        double tau;
        {
CDC_NO_CONTRACT
            tau = y * g[1] - x * g[2];
        }
        const size_t ti = (size_t)b * 2 + k;
        const double s1 = tq_c + tau;
        A.tq_o[ti] = A.read_base ? s1 * A.c1 + tq_b * A.c0 : s1;
    }
    if (lane < 7 * 4) {
        if (lane < 7 * NV) {
            const double s1 = ist_c + s_ds[lane];
            A.ist_o[oi + lane] = A.read_base ? s1 * A.c1 + ist_b * A.c0 : s1;
        }
        const double s1 = oec_c + s_de[lane];
        A.oec_o[oe + lane] = A.read_base ? s1 * A.c1 + oec_b * A.c0 : s1;
    }
}

}  // namespace cdc
