// cdc_orbit.h -- two-body orbit and its perturbation, for host and device.
//
// The reference takes these from the kepler-two-body crate (git dependency
// @2d8a0f5, not in the reference tree; call sites src/physics.rs:411-428,
// src/scheme.rs:214-215, src/state.rs:261-269).  The published algorithm is
// restated: Newton solve of Kepler's equation, pericentre on +x at t = 0,
// body 1 at -(m2/M) r_rel and body 2 at +(m1/M) r_rel with m1 = M/(1+q),
// m2 = M q/(1+q); the perturbation adds mass and momentum to each body at
// the orbital state of time t and recovers (a, M, q, e) from the relative
// orbit (energy -> a, eccentricity vector -> e).
//
// Both are compiled without floating-point contraction so that the host and
// device evaluations of perturb() round identically.
#pragma once

#include <math.h>

#include "../../include/cdc_solver.h"

#if defined(__HIPCC__)
#define CDC_HD __host__ __device__
#else
#define CDC_HD
#endif

// Host builds use gcc with -ffp-contract=off; clang needs the pragma.
#if defined(__clang__)
#define CDC_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define CDC_NO_CONTRACT
#endif

namespace cdc {

struct Orbit {
    double m1, x1, y1, vx1, vy1;
    double m2, x2, y2, vx2, vy2;
};

// Physics::orbital_elements, src/physics.rs:411-417
inline void orbital_elements(const cdc_physics_t &ph, double el[4]) {
    if (ph.one_body) {
        el[0] = 1e-12; el[1] = 1.0; el[2] = 1.0; el[3] = 0.0;
    } else {
        el[0] = 1.0; el[1] = 1.0; el[2] = ph.binary_mass_ratio; el[3] = ph.binary_eccentricity;
    }
}

// OrbitalElements::orbital_state_from_time (host only: libm sin/cos)
inline Orbit orbital_state(const double el[4], double t) {
CDC_NO_CONTRACT
    const double a = el[0], M = el[1], q = el[2], e = el[3];
    const double m1 = M / (1.0 + q);
    const double m2 = M * q / (1.0 + q);
    const double n = sqrt(M / (a * a * a));
    const double mean = n * t;
    double E = mean;
    for (int k = 0; k < 64; ++k) {
        const double f = E - e * sin(E) - mean;
        const double fp = 1.0 - e * cos(E);
        const double dE = f / fp;
        E = E - dE;
        if (fabs(dE) <= 1e-15 * (1.0 + fabs(E))) break;
    }
    const double cE = cos(E), sE = sin(E);
    const double b = sqrt(1.0 - e * e);
    const double x = a * (cE - e);
    const double y = a * b * sE;
    const double den = 1.0 - e * cE;
    const double vx = -a * n * sE / den;
    const double vy = a * n * b * cE / den;
    const double f1 = m2 / M, f2 = m1 / M;
    Orbit o;
    o.m1 = m1; o.x1 = -f1 * x; o.y1 = -f1 * y; o.vx1 = -f1 * vx; o.vy1 = -f1 * vy;
    o.m2 = m2; o.x2 = f2 * x;  o.y2 = f2 * y;  o.vx2 = f2 * vx;  o.vy2 = f2 * vy;
    return o;
}

// OrbitalElements::perturb(t, dm1, dm2, dpx1, dpx2, dpy1, dpy2) at the
// orbital state `s` of time t.  Writes the new elements; returns false when
// the perturbed binary is unbound (kepler_two_body::UnboundOrbitalState).
CDC_HD inline bool perturb(const Orbit &s, double dm1, double dm2, double dpx1, double dpx2,
                           double dpy1, double dpy2, double out[4]) {
CDC_NO_CONTRACT
    const double m1 = s.m1 + dm1;
    const double m2 = s.m2 + dm2;
    const double vx1 = (s.m1 * s.vx1 + dpx1) / m1;
    const double vy1 = (s.m1 * s.vy1 + dpy1) / m1;
    const double vx2 = (s.m2 * s.vx2 + dpx2) / m2;
    const double vy2 = (s.m2 * s.vy2 + dpy2) / m2;
    const double M = m1 + m2;
    const double q = m2 / m1;
    const double rx = s.x2 - s.x1;
    const double ry = s.y2 - s.y1;
    const double vx = vx2 - vx1;
    const double vy = vy2 - vy1;
    const double r = sqrt(rx * rx + ry * ry);
    const double energy = 0.5 * (vx * vx + vy * vy) - M / r;
    if (!(energy < 0.0) || !(m1 > 0.0) || !(m2 > 0.0)) return false;
    const double h = rx * vy - ry * vx;
    const double ex = vy * h / M - rx / r;
    const double ey = -vx * h / M - ry / r;
    out[0] = -M / (2.0 * energy);
    out[1] = M;
    out[2] = q;
    out[3] = sqrt(ex * ex + ey * ey);
    return true;
}

// ItemizedChange::perturbation, src/state.rs:258-282.  ds: [7][nv] block
// sums already scaled by dx*dy; de: [7][4].  Returns false on an unbound
// orbit.
CDC_HD inline bool perturbation(const Orbit &s, const double el[4], const double *ds, int nv,
                                double *de) {
CDC_NO_CONTRACT
    for (int k = 0; k < 7 * 4; ++k) de[k] = 0.0;
    double p[4];
    // sink1 -> body 1, sink2 -> body 2, grav1 -> body 1, grav2 -> body 2;
    // each body receives (-dm, -dp) of the gas change.
    const int items[4] = {CDC_ITEM_SINK1, CDC_ITEM_SINK2, CDC_ITEM_GRAV1, CDC_ITEM_GRAV2};
    for (int k = 0; k < 4; ++k) {
        const double *d = ds + items[k] * nv;
        bool ok = (k % 2 == 0) ? perturb(s, -d[0], 0.0, -d[1], 0.0, -d[2], 0.0, p)
                               : perturb(s, 0.0, -d[0], 0.0, -d[1], 0.0, -d[2], p);
        if (!ok) return false;
        for (int c = 0; c < 4; ++c) de[items[k] * 4 + c] = p[c] - el[c];
    }
    return true;
}

}  // namespace cdc
