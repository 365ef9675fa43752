"""Pins of the CPU oracle (oracle/cdc_oracle.c) before it is trusted as the
parity reference.  The reference has no tests, fixtures or golden vectors, and
its arithmetic crates are absent, so each building block is pinned by what
the mathematics fixes (closed forms, consistency, symmetry, conservation) or
by the reference's own independent restatement in
tools/indep_res_analysis.py, chosen so that a dropped term, a wrong sign, a
wrong index or a transposed operand fails at least one test.
"""
import math

import numpy as np
import pytest

from cdc_amd import Euler, Isothermal, Mesh, workloads
from oracle import oracle as O
from tests import helpers as H
from tests.numpy_reference import Step, to_blocks, to_global

O.load()


def phys_c(**kw):
    return workloads.base_physics(**kw).to_c()


# ---- HLLE (hydro-iso2d / hydro-euler riemann_hlle) ---------------------------

def iso_flux(p, axis, cs2):
    d, vx, vy = p
    vn = (vx, vy)[axis]
    f = np.array([d * vn, d * vx * vn, d * vy * vn])
    f[1 + axis] += d * cs2
    return f


def euler_flux(p, axis, gm):
    d, vx, vy, pg = p
    vn = (vx, vy)[axis]
    E = pg / (gm - 1) + 0.5 * d * (vx * vx + vy * vy)
    f = np.array([d * vn, d * vx * vn, d * vy * vn, (E + pg) * vn])
    f[1 + axis] += pg
    return f


@pytest.mark.parametrize("axis", [0, 1])
def test_hlle_iso_consistency_and_upwinding(axis):
    cs2 = 0.04
    rng = np.random.default_rng(1)
    for _ in range(20):
        p = np.array([rng.uniform(0.1, 2), rng.uniform(-0.1, 0.1), rng.uniform(-0.1, 0.1)])
        np.testing.assert_allclose(O.riemann_hlle(0, p, p, axis, cs2)[:3], iso_flux(p, axis, cs2),
                                   rtol=1e-14, atol=1e-15)
    # supersonic to the right: the flux is the left state's; to the left: the right's
    pl = np.array([1.0, 0.0, 0.0]); pr = np.array([0.5, 0.0, 0.0])
    pl[1 + axis] = 1.0; pr[1 + axis] = 0.9          # v_n - cs > 0 for both
    np.testing.assert_allclose(O.riemann_hlle(0, pl, pr, axis, cs2)[:3], iso_flux(pl, axis, cs2), rtol=1e-14)
    pl[1 + axis] = -1.0; pr[1 + axis] = -0.9
    np.testing.assert_allclose(O.riemann_hlle(0, pl, pr, axis, cs2)[:3], iso_flux(pr, axis, cs2), rtol=1e-14)


def test_hlle_iso_static_contact_value():
    # two static states: a+ = cs, a- = -cs, F = (F_L + F_R)/2 + cs (U_L - U_R)/2
    cs2 = 0.09
    cs = 0.3
    pl, pr = np.array([2.0, 0.0, 0.0]), np.array([1.0, 0.0, 0.0])
    f = O.riemann_hlle(0, pl, pr, 0, cs2)[:3]
    expect = 0.5 * (iso_flux(pl, 0, cs2) + iso_flux(pr, 0, cs2)) + 0.5 * cs * (pl - pr) * np.array([1, 0, 0])
    np.testing.assert_allclose(f, expect, rtol=1e-14, atol=1e-16)


@pytest.mark.parametrize("axis", [0, 1])
def test_hlle_euler_consistency_and_upwinding(axis):
    gm = 5.0 / 3.0
    rng = np.random.default_rng(2)
    for _ in range(20):
        p = np.array([rng.uniform(0.1, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.01, 1)])
        np.testing.assert_allclose(O.riemann_hlle(1, p, p, axis, gm), euler_flux(p, axis, gm), rtol=1e-13,
                                   atol=1e-15)
    pl = np.array([1.0, 0.0, 0.0, 0.01]); pr = np.array([0.8, 0.0, 0.0, 0.02])
    pl[1 + axis] = 3.0; pr[1 + axis] = 2.5
    np.testing.assert_allclose(O.riemann_hlle(1, pl, pr, axis, gm), euler_flux(pl, axis, gm), rtol=1e-14)


def test_hlle_mirror_symmetry():
    # reflecting x -> -x maps F(pl, pr) to -F(mirror pr, mirror pl) in the normal components
    gm = 1.4
    pl = np.array([1.0, 0.3, 0.1, 0.5]); pr = np.array([0.7, -0.2, 0.05, 0.3])
    f = O.riemann_hlle(1, pl, pr, 0, gm)
    ml = pr * np.array([1, -1, 1, 1]); mr = pl * np.array([1, -1, 1, 1])
    g = O.riemann_hlle(1, ml, mr, 0, gm)
    np.testing.assert_allclose(f * np.array([-1, 1, -1, -1]), g, rtol=1e-14, atol=1e-16)


# ---- PLM (godunov-core plm_gradient) ------------------------------------------

def test_plm_cases():
    assert O.plm_gradient(1.5, 0.0, 1.0, 3.0) == 1.5          # theta-limited left difference
    assert O.plm_gradient(2.0, 0.0, 1.0, 2.0) == 1.0          # central difference
    assert O.plm_gradient(1.5, 0.0, 2.0, 2.5) == 0.75         # theta-limited right difference
    assert O.plm_gradient(1.5, 0.0, 1.0, 0.5) == 0.0          # extremum
    assert O.plm_gradient(0.0, 0.0, 1.0, 3.0) == 0.0          # theta = 0: first order
    assert O.plm_gradient(1.5, 3.0, 1.0, 0.0) == -1.5         # decreasing data
    rng = np.random.default_rng(3)
    for _ in range(200):
        a, b, c = rng.normal(size=3)
        g = O.plm_gradient(1.7, a, b, c)
        # TVD bound and sign agreement with the central difference
        assert abs(g) <= max(1.7 * abs(b - a), 0) + 1e-15 and abs(g) <= 1.7 * abs(c - b) + 1e-15
        assert g == 0.0 or np.sign(g) == np.sign(c - a)


# ---- Plummer gravity (kepler-two-body) ----------------------------------------

def test_plummer_acceleration_is_minus_grad_phi():
    st = np.array([[0.6, -0.3, 0.1, 0.0, -0.4], [0.4, 0.45, -0.15, 0.0, 0.6]])
    rs = 0.05
    h = 1e-6
    for (x, y) in [(0.3, 0.7), (-1.2, 0.4), (2.0, -3.0), (0.44, -0.12)]:
        phi, acc = O.gravity(st, x, y, rs)
        ax = acc[0] + acc[2]
        ay = acc[1] + acc[3]
        gx = -(O.gravity(st, x + h, y, rs)[0] - O.gravity(st, x - h, y, rs)[0]) / (2 * h)
        gy = -(O.gravity(st, x, y + h, rs)[0] - O.gravity(st, x, y - h, rs)[0]) / (2 * h)
        np.testing.assert_allclose([ax, ay], [gx, gy], rtol=1e-6)
        # tools/indep_res_analysis.py:182-194 form
        for k in range(2):
            m, bx, by = st[k, 0], st[k, 1], st[k, 2]
            pre = -m / ((x - bx) ** 2 + (y - by) ** 2 + rs ** 2) ** 1.5
            np.testing.assert_allclose(acc[2 * k:2 * k + 2], [pre * (x - bx), pre * (y - by)], rtol=1e-14)


# ---- viscous stress (CellData::stress_field) vs tools/indep_res_analysis.py ----

def test_stress_field_matches_independent_form():
    rng = np.random.default_rng(4)
    for _ in range(20):
        pc = np.array([rng.uniform(0.5, 2), 0, 0, 0])
        gx = np.concatenate([[0], rng.normal(size=3)])
        gy = np.concatenate([[0], rng.normal(size=3)])
        nu, lm, dx, dy = 1e-3, 0.7, 0.1, 0.1
        lam = lm * nu
        dvx_dx, dvy_dy, dvy_dx, dvx_dy = gx[1] / dx, gy[2] / dy, gx[2] / dx, gy[1] / dy
        div = dvx_dx + dvy_dy
        txx = pc[0] * nu * (2 * dvx_dx - 2.0 / 3 * div) + pc[0] * lam * div
        tyy = pc[0] * nu * (2 * dvy_dy - 2.0 / 3 * div) + pc[0] * lam * div
        txy = pc[0] * nu * (dvx_dy + dvy_dx)
        got = [O.stress_field(pc, gx, gy, nu, lam, dx, dy, r, c) for r, c in [(0, 0), (0, 1), (1, 0), (1, 1)]]
        np.testing.assert_allclose(got, [txx, txy, txy, tyy], rtol=1e-12, atol=1e-15)


# ---- sinks, buffer, cooling, gravity sources ------------------------------------

def test_sink_kernel_values():
    p = phys_c(sink_radius=0.2, sink_rate=7.0)
    assert O.sink_kernel(p, 0.0, 0.0) == 7.0
    assert O.sink_kernel(p, 0.1, 0.0) == pytest.approx(7.0 * 0.75 ** 2, rel=1e-15)
    assert O.sink_kernel(p, 0.0, 0.2) == 0.0
    assert O.sink_kernel(p, 0.3, 0.3) == 0.0


def _one_state(t=0.3, **kw):
    return O.orbital_state(phys_c(**kw), t)


def test_iso_sources_against_independent_forms():
    mesh = Mesh(4, 16, 8.0).to_c()
    hy = Isothermal(mach_number=10.0).to_c()
    ph = workloads.base_physics(sink_radius=0.3, buffer_scale=1.0)
    pc = ph.to_c()
    st = _one_state()
    rs = ph.softening()
    dt = 1e-3
    for (x, y) in [(0.5, 0.1), (6.9, 3.3), (-7.5, -7.9), (st[0, 1] + 0.1, st[0, 2])]:
        u = np.array([0.7, 0.2, -0.1])
        u0 = np.array([1.0, 0.0, 0.3])
        it = O.source_terms(mesh, hy, pc, u, u0, x, y, 0.3, dt, st)
        # gravity (indep_res_analysis.py:182-194)
        for k in range(2):
            m, bx, by = st[k, :3]
            pre = -u[0] * m / ((x - bx) ** 2 + (y - by) ** 2 + rs ** 2) ** 1.5
            np.testing.assert_allclose(it[2 + k, :3], [0, pre * (x - bx) * dt, pre * (y - by) * dt], rtol=1e-13)
            r2 = (x - bx) ** 2 + (y - by) ** 2
            rate = 5.0 * 0 + (ph.sink_rate * (1 - r2 / ph.sink_radius ** 2) ** 2 if r2 < ph.sink_radius ** 2 else 0)
            np.testing.assert_allclose(it[k, :3], -rate * u * dt, rtol=1e-13, atol=1e-300)
        # buffer (indep_res_analysis.py:196-213)
        R = 8.0
        w = 0.5 * ph.buffer_rate * (1 + np.tanh((np.hypot(x, y) - R) / ph.buffer_scale)) * np.sqrt(1 / R ** 3)
        np.testing.assert_allclose(it[4, :3], -w * (u - u0) * dt, rtol=1e-13)
        assert np.all(it[5] == 0)


def test_euler_t4_cooling_is_exact_solution_of_de_dt():
    mesh = Mesh(4, 16, 8.0).to_c()
    gm = 5.0 / 3.0
    hyd = Euler(gamma_law_index=gm, cooling_coefficient=0.5, density_index=2)
    ph = workloads.base_physics(sink_radius=0.01, buffer_rate=0.0)
    st = _one_state()
    d, vx, vy, p = 0.8, 0.1, -0.2, 0.3
    e0 = p / d / (gm - 1)
    u = np.array([d, d * vx, d * vy, p / (gm - 1) + 0.5 * d * (vx * vx + vy * vy)])
    a = hyd.cooling_coefficient / d ** 2

    def cooled(dt):
        it = O.source_terms(mesh, hyd.to_c(), ph.to_c(), u, u, 3.0, 2.0, 0.0, dt, st)
        return (u[3] + it[5, 3] - 0.5 * d * (vx * vx + vy * vy)) / d

    # integrate de/dt = -a e^4 with classical RK4 in small steps
    for dt in (1e-3, 0.1, 2.0):
        e, n = e0, 4000
        h = dt / n
        for _ in range(n):
            k1 = -a * e ** 4; k2 = -a * (e + h / 2 * k1) ** 4
            k3 = -a * (e + h / 2 * k2) ** 4; k4 = -a * (e + h * k3) ** 4
            e += h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        assert cooled(dt) == pytest.approx(e, rel=1e-10)
    # small dt: energy rate -coef e^4 / Sigma (indep_res_analysis.py:215-216)
    dt = 1e-9
    it = O.source_terms(mesh, hyd.to_c(), ph.to_c(), u, u, 3.0, 2.0, 0.0, dt, st)
    assert it[5, 3] / dt == pytest.approx(-hyd.cooling_coefficient * e0 ** 4 / d, rel=1e-5)
    assert abs(it[5, 0]) == 0.0


def test_euler_beta_cooling_relaxes_exponentially():
    mesh = Mesh(4, 16, 8.0).to_c()
    gm = 5.0 / 3.0
    hyd = Euler(gamma_law_index=gm, cooling_coefficient=0.0, cooling_beta=10.0)
    ph = workloads.base_physics(sink_radius=0.01, buffer_rate=0.0)
    st = _one_state()
    x, y, dt = 2.0, 1.0, 0.37
    d = 0.9
    u = np.array([d, 0.0, 0.0, 0.5 / (gm - 1)])
    bg = np.array([1.0, 0.0, 0.0, 0.2 / (gm - 1)])
    it = O.source_terms(mesh, hyd.to_c(), ph.to_c(), u, bg, x, y, 0.0, dt, st)
    e0, et = 0.5 / d / (gm - 1), 0.2 / 1.0 / (gm - 1)
    t2 = x * x + y * y + ph.softening() ** 2
    om = math.sqrt(1.0 / t2 ** 1.5)
    expect = et + (e0 - et) * math.exp(-dt * om / 10.0)
    assert (u[3] + it[5, 3]) / d == pytest.approx(expect, rel=1e-13)


# ---- orbit (kepler-two-body) ----------------------------------------------------

def test_circular_orbit_closed_form():
    p = phys_c(binary_eccentricity=0.0, binary_mass_ratio=1.0)
    for t in [0.0, 0.7, math.pi / 2, 2 * math.pi, 13.1]:
        s = O.orbital_state(p, t)
        np.testing.assert_allclose(s[1, 1:3] - s[0, 1:3], [math.cos(t), math.sin(t)], atol=1e-14)
        np.testing.assert_allclose(s[0, 0] * s[0, 1:] + s[1, 0] * s[1, 1:], 0.0, atol=1e-15)  # COM at rest
    s0, s1 = O.orbital_state(p, 0.0), O.orbital_state(p, 2 * math.pi)
    np.testing.assert_allclose(s0, s1, atol=1e-14)


@pytest.mark.parametrize("e,q", [(0.5, 0.5), (0.8, 0.1), (0.2, 1.0)])
def test_eccentric_orbit_invariants(e, q):
    p = phys_c(binary_eccentricity=e, binary_mass_ratio=q)
    M, a = 1.0, 1.0
    for t in [0.0, 0.3, 1.9, math.pi, 5.5]:
        s = O.orbital_state(p, t)
        r = s[1, 1:3] - s[0, 1:3]
        v = s[1, 3:5] - s[0, 3:5]
        rn = np.hypot(*r)
        assert 0.5 * v @ v - M / rn == pytest.approx(-M / (2 * a), rel=1e-13)
        assert r[0] * v[1] - r[1] * v[0] == pytest.approx(math.sqrt(M * a * (1 - e * e)), rel=1e-13)
        assert s[1, 0] / s[0, 0] == pytest.approx(q, rel=1e-15)
        np.testing.assert_allclose(s[0, 0] * s[0, 1:] + s[1, 0] * s[1, 1:], 0.0, atol=1e-15)
        h = 1e-6
        dr = (O.orbital_state(p, t + h)[1, 1:3] - O.orbital_state(p, t - h)[1, 1:3]) / (2 * h)
        np.testing.assert_allclose(dr, s[1, 3:5], rtol=1e-7, atol=1e-8)
    assert np.hypot(*(O.orbital_state(p, 0.0)[1, 1:3] - O.orbital_state(p, 0.0)[0, 1:3])) == pytest.approx(1 - e)
    sh = O.orbital_state(p, math.pi)
    assert np.hypot(*(sh[1, 1:3] - sh[0, 1:3])) == pytest.approx(1 + e, rel=1e-13)


def test_perturb_closed_forms():
    p = phys_c(binary_eccentricity=0.0, binary_mass_ratio=1.0)
    el = O.orbital_elements(p)
    st = O.orbital_state(p, 0.4)
    rc, de = O.perturb(st, el, np.zeros(6))
    assert rc == 0 and np.max(np.abs(de)) < 1e-14
    # mass added to body 1 at rest relative to it: M and q change exactly
    dm = 1e-3
    rc, de = O.perturb(st, el, [dm, 0, dm * st[0, 3], 0, dm * st[0, 4], 0])
    assert rc == 0
    assert de[1] == pytest.approx(dm, rel=1e-12)
    assert de[2] == pytest.approx(0.5 / (0.5 + dm) - 1.0, rel=1e-10)
    # tangential kick: vis-viva a' = 1 / (2/r - v'^2/M)
    kick = 1e-3
    v2 = st[1, 3:5] * (1 + kick / (0.5 * np.hypot(*st[1, 3:5])))
    dp = 0.5 * (v2 - st[1, 3:5])
    rc, de = O.perturb(st, el, [0, 0, 0, dp[0], 0, dp[1]])
    vrel = np.hypot(*(v2 - st[0, 3:5]))
    assert de[0] == pytest.approx(1.0 / (2.0 / 1.0 - vrel ** 2) - 1.0, rel=1e-9)
    assert de[3] > 0
    # a large kick unbinds the binary
    rc, _ = O.perturb(st, el, [0, 0, 0, 10 * st[1, 3], 0, 10 * st[1, 4]])
    assert rc == 3


# ---- geometry -------------------------------------------------------------------

def test_block_vertices_tile_the_domain():
    m = Mesh(5, 12, 7.0)
    mc = m.to_c()
    L = 2 * 7.0 / 5
    for b in range(5):
        xv = O.block_vertices(mc, b)
        assert xv[0] == -7.0 + b * L
        np.testing.assert_allclose(np.diff(xv), L / 12, rtol=1e-12)
        np.testing.assert_array_equal(xv, m.block_vertices((b, 0))[0])


# ---- whole-step invariants and the brute-force restatement ------------------------

def run_oracle(mesh, hydro, physics, u0, steps, fold=1):
    o = H.oracle_for(mesh, hydro, physics, u0)
    dts = [o.advance(fold) for _ in range(steps)]
    return o, dts


@pytest.mark.parametrize("kind,visc,rk,sink", [
    ("iso", {"constant_nu": 1e-3}, "RK2", "impartial"),
    ("iso", {"alpha": 0.1}, "RK3", "torque_free"),
    ("euler", {"alpha": 0.1}, "RK3", "impartial"),
    ("euler_beta", {"constant_nu": 2e-3}, "RK1", "torque_free"),
])
def test_oracle_matches_numpy_global_restatement(kind, visc, rk, sink):
    mesh, hydro, ph = H.small(kind, nb=2, bs=8, R=4.0, viscosity=visc, rk_order=rk, sink_model=sink,
                              binary_eccentricity=0.3, binary_mass_ratio=0.6, lambda_multiplier=0.5)
    u0 = H.initial(mesh, hydro, ph, cavity=1.5)
    o, _ = run_oracle(mesh, hydro, ph, u0, 0)
    ref = Step(mesh, hydro, ph)
    U0g = to_global(u0, 2, 8)
    u = U0g.copy()
    ist = np.zeros((4, 7, hydro.nv))
    t = 0.0
    for _ in range(3):
        dt = o.advance(1)
        u, ist, t = ref.step(u, U0g, ist, t, dt)
        s = o.get_state()
        assert s["time"] == pytest.approx(t, rel=1e-14)
        assert H.field_err(s["conserved"], to_blocks(u, 2, 8)) < 1e-12
        # cooling = to_conserved(cooled primitive) - u is a difference of nearly
        # equal energies (condition number E / dE), so it gets a looser bound
        keep = [0, 1, 2, 3, 4, 6]
        assert H.items_err(s["ist"][:, keep], ist[:, keep]) < 1e-10
        if hydro.nv == 4:
            ce, cr = s["ist"][:, 5, 3], ist[:, 5, 3]
            assert np.max(np.abs(ce - cr)) <= 1e-6 * np.max(np.abs(cr))


@pytest.mark.parametrize("kind", ["iso", "euler"])
def test_block_decomposition_invariance(kind):
    states = []
    for nb, bs in [(4, 8), (2, 16), (1, 32)]:
        mesh, hydro, ph = H.small(kind, nb=nb, bs=bs, R=4.0, rk_order="RK2")
        u0 = H.initial(mesh, hydro, ph)
        o, _ = run_oracle(mesh, hydro, ph, u0, 3)
        states.append(to_global(o.get_state()["conserved"], nb, bs))
    for s in states[1:]:
        assert H.field_err(s, states[0]) < 1e-13


@pytest.mark.parametrize("kind,rk", [("iso", "RK2"), ("iso", "RK3"), ("euler", "RK3"), ("euler", "RK1")])
def test_mass_and_momentum_budget_closure(kind, rk):
    """Fluxes telescope on the periodic mesh: the change of total mass and
    momentum equals the integrated source terms (src/scheme.rs:307-313).
    Each block evaluates its own boundary faces (src/scheme.rs:257-268), so
    the isothermal flux at the periodic edge sees cs^2 = -phi/M^2 at x = -R
    from one block and at x = +R from the other; the two agree only when
    the potential is symmetric, so the isothermal case uses one_body."""
    mesh, hydro, ph = H.small(kind, nb=2, bs=16, R=4.0, rk_order=rk, fake_mass_rate=0.0,
                              one_body=(kind == "iso"))
    u0 = H.initial(mesh, hydro, ph)
    o, _ = run_oracle(mesh, hydro, ph, u0, 4)
    s = o.get_state()
    dA = mesh.cell_spacing() ** 2
    du = (s["conserved"].sum(axis=(0, 1, 2)) - u0.sum(axis=(0, 1, 2))) * dA
    src = s["ist"].sum(axis=(0, 1))
    scale = np.abs(u0).sum(axis=(0, 1, 2)) * dA
    np.testing.assert_allclose(du[:3], src[:3], atol=1e-12 * scale.max())
    # mass sources are sinks, buffer and fake mass only
    assert np.all(s["ist"][:, 2:4, 0] == 0.0)


def test_point_symmetry_equal_mass_circular():
    mesh, hydro, ph = H.small("iso", nb=2, bs=16, R=4.0)
    u0 = H.initial(mesh, hydro, ph, noise=0.0)
    o, _ = run_oracle(mesh, hydro, ph, u0, 5)
    u = to_global(o.get_state()["conserved"], 2, 16)
    rot = u[::-1, ::-1] * np.array([1, -1, -1])
    assert H.field_err(rot, u) < 1e-12


def test_quarter_turn_symmetry_single_body():
    """one_body: the setup is invariant under a 90 degree rotation, which
    catches x/y transpositions in gradients, faces and stresses."""
    mesh, hydro, ph = H.small("iso", nb=2, bs=16, R=4.0, one_body=True, viscosity={"constant_nu": 5e-3})
    u0 = H.initial(mesh, hydro, ph, noise=0.0)
    o, _ = run_oracle(mesh, hydro, ph, u0, 5)
    u = to_global(o.get_state()["conserved"], 2, 16)
    # (x, y) -> (-y, x): new[i, j] = old[j, N-1-i], velocity (vx, vy) -> (-vy, vx)
    N = u.shape[0]
    rot = np.empty_like(u)
    for i in range(N):
        for j in range(N):
            a = u[j, N - 1 - i]
            rot[i, j] = [a[0], -a[2], a[1]]
    assert H.field_err(rot, u) < 1e-12


def test_rk_time_and_iteration_bookkeeping():
    for rk in ("RK1", "RK2", "RK3"):
        mesh, hydro, ph = H.small("iso", nb=1, bs=16, R=4.0, rk_order=rk)
        u0 = H.initial(mesh, hydro, ph)
        o, dts = run_oracle(mesh, hydro, ph, u0, 3, fold=2)
        s = o.get_state()
        assert s["iteration"] == (6, 1)
        assert s["time"] == pytest.approx(2 * sum(dts), rel=1e-14)
        assert dts[0] == pytest.approx(ph.cfl * mesh.cell_spacing() / max(
            np.max(np.abs(u0[..., 1] / u0[..., 0])), np.max(np.abs(u0[..., 2] / u0[..., 0]))), rel=1e-15)


def test_negative_density_error_is_transactional():
    from cdc_amd import abi
    mesh, hydro, ph = H.small("iso", nb=3, bs=8, R=4.0)
    u0 = H.initial(mesh, hydro, ph)
    bad = u0.copy()
    bad[4, 2, 5, 0] = -1e-3       # block (1, 1), cell (2, 5)
    bad[8, 7, 7, 0] = -2e-3       # block (2, 2), cell (7, 7)
    o = H.oracle_for(mesh, hydro, ph, u0)
    o.set_state(bad, time=0.5, iteration=(3, 1))
    with pytest.raises(O.OracleError) as ei:
        o.advance(1)
    err = abi.CError.from_buffer_copy(ei.value.err[:__import__("ctypes").sizeof(abi.CError)])
    assert err.code == abi.CDC_ERR_NEGATIVE_DENSITY
    # block (0, 0) is the first to fail; with periodic wrap its first 3x3
    # neighbour in try_join_3by3 order is (nb-1, nb-1) = (2, 2), which it
    # reaches before (1, 1)
    assert err.value == -2e-3
    xc, yc = mesh.cell_centers((2, 2))
    assert (err.x, err.y) == (xc[7, 7], yc[7, 7])
    s = o.get_state()
    assert s["time"] == 0.5 and s["iteration"] == (3, 1)
    np.testing.assert_array_equal(s["conserved"], bad)


# ---- torque impulses (extension): central-force identity ------------------------

def test_torque_is_minus_the_sum_of_cell_torques():
    """Body k's torque impulse x_k x (-F_k) equals minus the sum over cells of
    x x f_k, because each f_k points along x - x_k: checked against a
    per-cell reduction of the Plummer force on one RK1 step."""
    from tests import helpers as H
    mesh, hydro, ph = H.small("iso", nb=2, bs=8, R=4.0, rk_order="RK1", binary_eccentricity=0.4,
                              binary_mass_ratio=0.6, buffer_rate=0.0, sink_rate=0.0)
    u0 = H.initial(mesh, hydro, ph)
    o = H.oracle_for(mesh, hydro, ph, u0)
    dt = o.advance(1)
    tq = o.get_torque()
    orb = O.orbital_state(ph.to_c(), 0.0)
    mc = mesh.to_c()
    rs = ph.softening()
    nb, bs = mesh.num_blocks, mesh.block_size
    dA = mesh.cell_spacing() ** 2
    for b in range(nb * nb):
        xv, yv = O.block_vertices(mc, b // nb), O.block_vertices(mc, b % nb)
        xc, yc = 0.5 * (xv[:-1] + xv[1:]), 0.5 * (yv[:-1] + yv[1:])
        want = np.zeros(2)
        mag = np.zeros(2)
        for i in range(bs):
            for j in range(bs):
                sig = u0[b, i, j, 0]
                _, acc = O.gravity(orb, xc[i], yc[j], rs)
                for k in range(2):
                    fx, fy = sig * acc[2 * k] * dt * dA, sig * acc[2 * k + 1] * dt * dA
                    want[k] -= xc[i] * fy - yc[j] * fx
                    mag[k] += abs(xc[i] * fy) + abs(yc[j] * fx)
        np.testing.assert_allclose(tq[b], want, rtol=0, atol=1e-12 * mag.max())


# ---- max signal speed, perturb's four outputs, RK weights: closed forms ----------

def _uniform_oracle(kind, prim, nb=1, bs=8, R=4.0, **phys):
    """An oracle whose whole mesh holds one primitive state."""
    mesh, hydro, ph = H.small(kind, nb=nb, bs=bs, R=R, **phys)
    nv = hydro.nv
    d, vx, vy = prim[0], prim[1], prim[2]
    u = [d, d * vx, d * vy]
    if nv == 4:
        u.append(prim[3] / (hydro.gamma_law_index - 1.0) + 0.5 * d * (vx * vx + vy * vy))
    u0 = np.broadcast_to(np.array(u), (nb * nb, bs, bs, nv)).copy()
    return H.oracle_for(mesh, hydro, ph, u0), hydro


@pytest.mark.parametrize("v", [(0.0, 0.0), (0.7, 0.0), (0.0, -0.7), (0.5, -0.5), (-0.3, 1.1)])
def test_euler_max_signal_speed_closed_form(v):
    """Euler Primitive::max_signal_speed (src/physics.rs:679-681, hydro-euler):
    the larger |v| component plus c_s = sqrt(gamma p / Sigma).  Dropping gamma,
    using |v| instead of the larger component, or leaving out c_s all fail."""
    d, p = 2.0, 0.3
    o, hydro = _uniform_oracle("euler", (d, v[0], v[1], p))
    gm = hydro.gamma_law_index
    expect = max(abs(v[0]), abs(v[1])) + math.sqrt(gm * p / d)
    assert o.max_signal_speed() == pytest.approx(expect, rel=1e-14)
    assert abs(expect - (max(abs(v[0]), abs(v[1])) + math.sqrt(p / d))) > 1e-3      # gamma matters
    if v[0] and v[1]:
        assert abs(expect - (math.hypot(*v) + math.sqrt(gm * p / d))) > 1e-3       # not |v|


@pytest.mark.parametrize("v", [(0.0, 0.0), (0.7, 0.0), (0.5, -0.9)])
def test_iso_max_signal_speed_has_no_sound_speed(v):
    """Isothermal max_signal_speed (src/physics.rs:521-526): max(|vx|, |vy|)
    with no sound speed, as written in the reference."""
    o, _ = _uniform_oracle("iso", (1.5, v[0], v[1]))
    assert o.max_signal_speed() == pytest.approx(max(abs(v[0]), abs(v[1])), rel=1e-15)


def _pert(st, el, dm=(0.0, 0.0), dp1=(0.0, 0.0), dp2=(0.0, 0.0)):
    """oracle perturb with delta = (dm1, dm2, dpx1, dpx2, dpy1, dpy2)."""
    return O.perturb(st, el, [dm[0], dm[1], dp1[0], dp2[0], dp1[1], dp2[1]])


def _kick_to_relative(st, dv):
    """Momentum kick on body 2 that changes the relative velocity by dv."""
    return st[1, 0] * np.asarray(dv)


def test_perturb_tangential_kick_of_circular_binary():
    """A circular binary (r = a = 1, M = 1) whose relative speed becomes v' along
    the orbit: r is an apsis, so e' = |r v'^2 / M - 1|, a' = 1 / (2/r - v'^2/M)
    (vis-viva), and M, q are unchanged."""
    p = phys_c(binary_eccentricity=0.0, binary_mass_ratio=0.5)
    el = O.orbital_elements(p)
    st = O.orbital_state(p, 0.9)
    vrel = st[1, 3:5] - st[0, 3:5]
    vc = np.hypot(*vrel)
    for f in (1.05, 0.9, 1.2):
        vp = f * vc
        rc, de = _pert(st, el, dp2=_kick_to_relative(st, (f - 1.0) * vrel))
        assert rc == 0
        assert de[3] == pytest.approx(abs(vp * vp - 1.0), rel=1e-10)
        assert de[0] == pytest.approx(1.0 / (2.0 - vp * vp) - 1.0, rel=1e-10)
        assert abs(de[1]) < 1e-15 and abs(de[2]) < 1e-15


def test_perturb_radial_kick_of_circular_binary():
    """A radial velocity v_r added to a circular binary: h is unchanged and
    e' = |v_r| / v_c exactly; a' = 1 / (2/r - (v_c^2 + v_r^2)/M)."""
    p = phys_c(binary_eccentricity=0.0, binary_mass_ratio=1.0)
    el = O.orbital_elements(p)
    st = O.orbital_state(p, 2.1)
    r = st[1, 1:3] - st[0, 1:3]
    rhat = r / np.hypot(*r)
    vc = np.hypot(*(st[1, 3:5] - st[0, 3:5]))
    for vr in (0.01, -0.05, 0.2):
        rc, de = _pert(st, el, dp2=_kick_to_relative(st, vr * rhat))
        assert rc == 0
        assert de[3] == pytest.approx(abs(vr) / vc, rel=1e-10)
        assert de[0] == pytest.approx(1.0 / (2.0 - (vc * vc + vr * vr)) - 1.0, rel=1e-10)


def test_perturb_pericentre_kick_of_eccentric_binary():
    """An e = 0.5 binary at pericentre (t = 0, r_p = a (1 - e) = 0.5) whose
    relative speed is raised to v' along the orbit: it stays the pericentre,
    so e' = r_p v'^2 / M - 1; a mass dm added at the body's velocity raises M
    and m1 only, so e' = r_p v_p^2 / (M + dm) - 1 and q' = m2 / (m1 + dm)."""
    e, q = 0.5, 0.5
    p = phys_c(binary_eccentricity=e, binary_mass_ratio=q)
    el = O.orbital_elements(p)
    st = O.orbital_state(p, 0.0)
    rp = np.hypot(*(st[1, 1:3] - st[0, 1:3]))
    assert rp == pytest.approx(1.0 - e, rel=1e-14)
    vrel = st[1, 3:5] - st[0, 3:5]
    vpn = np.hypot(*vrel)
    assert rp * vpn ** 2 - 1.0 == pytest.approx(e, rel=1e-12)
    for f in (1.02, 1.1):
        rc, de = _pert(st, el, dp2=_kick_to_relative(st, (f - 1.0) * vrel))
        assert rc == 0
        assert de[3] + e == pytest.approx(rp * (f * vpn) ** 2 - 1.0, rel=1e-11)
        assert de[0] + 1.0 == pytest.approx(1.0 / (2.0 / rp - (f * vpn) ** 2), rel=1e-11)
    dm = 2e-3
    m1, m2 = st[0, 0], st[1, 0]
    rc, de = _pert(st, el, dm=(dm, 0.0), dp1=dm * st[0, 3:5])
    assert rc == 0
    assert de[1] == pytest.approx(dm, rel=1e-12)
    assert de[2] + q == pytest.approx(m2 / (m1 + dm), rel=1e-13)
    assert de[3] + e == pytest.approx(rp * vpn ** 2 / (1.0 + dm) - 1.0, rel=1e-11)
    assert de[0] + 1.0 == pytest.approx(1.0 / (2.0 / rp - vpn ** 2 / (1.0 + dm)) * (1.0 + dm) / (1.0 + dm),
                                        rel=1e-11)


@pytest.mark.parametrize("rk,poly", [
    ("RK1", lambda z: 1 - z),
    ("RK2", lambda z: 1 - z + z * z / 2),
    ("RK3", lambda z: 1 - z + z * z / 2 - z ** 3 / 6),
])
def test_rk_weights_reproduce_taylor_polynomials(rk, poly):
    """SSP Runge-Kutta weights (godunov-core RungeKuttaOrder via
    WeightedAverage, src/scheme.rs:346, src/state.rs:183-200).  Gas at rest
    with uniform density, pressure and buffer rate (buffer_scale so large
    that tanh((r - R)/scale) = 0), no sinks or fake mass, and a softening
    length so large that the bodies' pull is below 1e-17: the state stays
    uniform and at rest, and the density obeys dSigma/dt = -k (Sigma -
    Sigma_0), so one step scales Sigma - Sigma_0 by the order-k Taylor
    polynomial of exp(-z), z = k dt.  Any other stage weights (b = 1/2 for
    RK2; 3/4, 1/3 for RK3) fail."""
    mesh, hydro, ph = H.small("euler", nb=1, bs=8, R=4.0, rk_order=rk, sink_rate=0.0, fake_mass_rate=0.0,
                              buffer_rate=40.0, buffer_scale=1e300, binary_mass_ratio=1.0,
                              binary_eccentricity=0.0, viscosity={"constant_nu": 0.0},
                              softening_length=1e6)
    hydro.cooling_coefficient = 0.0
    # a light disk, so that its pull on the binary stays far from unbinding it
    d0, p0 = 1e-6, 1e-6
    u_init = np.zeros((1, 8, 8, 4))
    u_init[..., 0] = d0
    u_init[..., 3] = p0 / (hydro.gamma_law_index - 1.0)
    o = H.oracle_for(mesh, hydro, ph, u_init)
    u = u_init.copy()
    d = 3e-6
    u[..., 0] = d
    u[..., 3] = p0 / (hydro.gamma_law_index - 1.0)
    o.set_state(u)
    dt = o.advance(1)
    k = 0.5 * ph.buffer_rate * math.sqrt(1.0 / mesh.domain_radius ** 3)
    z = k * dt
    assert 0.05 < z < 1.0
    got = (o.get_state()["conserved"][..., 0] - d0) / (d - d0)
    np.testing.assert_allclose(got, poly(z), rtol=1e-12)
    others = [1 - z, 1 - z + z * z / 2, 1 - z + z * z / 2 - z ** 3 / 6]
    assert sum(abs(poly(z) - w) > 1e-6 for w in others) == 2
