"""Independent numpy restatement of one RK step on the monolithic periodic grid.

Used only to pin the oracle's block/guard/stencil structure on tiny inputs:
where the oracle assembles each block's 2-zone guard from its 3x3 periodic
neighbours (src/scheme.rs:126-127) and walks faces per block
(src/scheme.rs:257-268), this module works on the whole N x N periodic array
with np.roll.  Formulas follow the textbook forms (tools/indep_res_analysis.py
writes the same viscous stress, gravity and buffer terms in the same style).
"""
import numpy as np


def kepler_state(el, t):
    a, M, q, e = el
    m1, m2 = M / (1 + q), M * q / (1 + q)
    n = np.sqrt(M / a ** 3)
    mean = n * t
    E = mean
    for _ in range(100):
        E = E - (E - e * np.sin(E) - mean) / (1 - e * np.cos(E))
    x, y = a * (np.cos(E) - e), a * np.sqrt(1 - e * e) * np.sin(E)
    den = 1 - e * np.cos(E)
    vx, vy = -a * n * np.sin(E) / den, a * n * np.sqrt(1 - e * e) * np.cos(E) / den
    return [(m1, -m2 / M * x, -m2 / M * y, -m2 / M * vx, -m2 / M * vy),
            (m2, m1 / M * x, m1 / M * y, m1 / M * vx, m1 / M * vy)]


def minmod(a, b, c):
    return np.where((a > 0) & (b > 0) & (c > 0), np.minimum(np.minimum(a, b), c),
                    np.where((a < 0) & (b < 0) & (c < 0), np.maximum(np.maximum(a, b), c), 0.0))


def global_coords(mesh):
    """Per cell: left vertex, right vertex and centre, each from the cell's own
    block (each block evaluates both of its boundary faces, src/scheme.rs:257-268,
    so the periodic faces at -R and +R see different coordinates)."""
    nb = mesh.num_blocks
    vs = [mesh.block_vertices((b, b))[0] for b in range(nb)]
    xl = np.concatenate([v[:-1] for v in vs])
    xr = np.concatenate([v[1:] for v in vs])
    xc = np.concatenate([(v[:-1] + v[1:]) / 2 for v in vs])
    return xl, xr, xc


def to_global(u_blocks, nb, bs):
    nv = u_blocks.shape[-1]
    return u_blocks.reshape(nb, nb, bs, bs, nv).transpose(0, 2, 1, 3, 4).reshape(nb * bs, nb * bs, nv)


def to_blocks(u, nb, bs):
    nv = u.shape[-1]
    return u.reshape(nb, bs, nb, bs, nv).transpose(0, 2, 1, 3, 4).reshape(nb * nb, bs, bs, nv)


class Step:
    def __init__(self, mesh, hydro, physics):
        self.mesh, self.hydro, self.ph = mesh, hydro, physics
        self.euler = hydro.nv == 4
        self.gm = hydro.gamma_law_index if self.euler else 1.0
        self.dx = mesh.cell_spacing()
        self.xl, self.xr, self.xc = global_coords(mesh)
        self.XC, self.YC = np.meshgrid(self.xc, self.xc, indexing="ij")
        self.el = (1e-12, 1.0, 1.0, 0.0) if physics.one_body else \
            (1.0, 1.0, physics.binary_mass_ratio, physics.binary_eccentricity)
        self.rs = physics.softening()

    def prim(self, u):
        d = u[..., 0]
        vx, vy = u[..., 1] / d, u[..., 2] / d
        if self.euler:
            p = (u[..., 3] - 0.5 * d * (vx ** 2 + vy ** 2)) * (self.gm - 1)
            return np.stack([d, vx, vy, p], -1)
        return np.stack([d, vx, vy], -1)

    def cons(self, P):
        d, vx, vy = P[..., 0], P[..., 1], P[..., 2]
        if self.euler:
            E = 0.5 * d * (vx ** 2 + vy ** 2) + P[..., 3] / (self.gm - 1)
            return np.stack([d, d * vx, d * vy, E], -1)
        return np.stack([d, d * vx, d * vy], -1)

    def phi(self, x, y, bodies):
        return sum(-m / np.sqrt((x - bx) ** 2 + (y - by) ** 2 + self.rs ** 2) for (m, bx, by, _, _) in bodies)

    def hlle(self, pl, pr, axis, cs2):
        ul, ur = self.cons(pl), self.cons(pr)
        vl, vr = pl[..., 1 + axis], pr[..., 1 + axis]
        if self.euler:
            cl, cr = np.sqrt(self.gm * pl[..., 3] / pl[..., 0]), np.sqrt(self.gm * pr[..., 3] / pr[..., 0])
            ppl, ppr = pl[..., 3], pr[..., 3]
        else:
            cl = cr = np.sqrt(cs2)
            ppl, ppr = pl[..., 0] * cs2, pr[..., 0] * cs2
        fl, fr = ul * vl[..., None], ur * vr[..., None]
        fl[..., 1 + axis] += ppl
        fr[..., 1 + axis] += ppr
        if self.euler:
            fl[..., 3] += ppl * vl
            fr[..., 3] += ppr * vr
        ap = np.maximum(np.maximum(vl + cl, vr + cr), 0)[..., None]
        am = np.minimum(np.minimum(vl - cl, vr - cr), 0)[..., None]
        return (fl * ap - fr * am - (ul - ur) * ap * am) / (ap - am)

    def stage(self, u, u_init, t, dt):
        """Returns (u1, ds[7][nv] per block) for one forward-Euler stage."""
        ph, dx = self.ph, self.dx
        bodies = kepler_state(self.el, t)
        P = self.prim(u)
        th = ph.plm
        g = [minmod(th * (P - np.roll(P, 1, ax)), 0.5 * (np.roll(P, -1, ax) - np.roll(P, 1, ax)),
                    th * (np.roll(P, -1, ax) - P)) for ax in (0, 1)]
        kind, val = ph.viscosity_model()
        lam = ph.lambda_multiplier
        fluxes = []
        for ax, side in ((0, 0), (0, 1), (1, 0), (1, 1)):
            # side 0: face a-1/2 of each cell (L = cell - 1, R = cell);
            # side 1: face a+1/2 (L = cell, R = cell + 1); positions from the cell's block
            sh = 1 if side == 0 else 0
            PLc, gLx, gLy = np.roll(P, sh, ax), np.roll(g[0], sh, ax), np.roll(g[1], sh, ax)
            PRc, gRx, gRy = np.roll(P, sh - 1, ax), np.roll(g[0], sh - 1, ax), np.roll(g[1], sh - 1, ax)
            gn_L, gn_R = (gLx, gRx) if ax == 0 else (gLy, gRy)
            pl, pr = PLc + 0.5 * gn_L, PRc - 0.5 * gn_R
            xf = self.xl if side == 0 else self.xr
            if ax == 0:
                fx, fy = np.meshgrid(xf, self.xc, indexing="ij")
            else:
                fx, fy = np.meshgrid(self.xc, xf, indexing="ij")
            ph_f = self.phi(fx, fy, bodies)
            cs2 = -ph_f / self.hydro.mach_number ** 2 if not self.euler else None

            def nu(Pc):
                if kind == "constant_nu":
                    return val
                (m1, x1, y1, _, _), (m2, x2, y2, _, _) = bodies
                r1 = np.sqrt((fx - x1) ** 2 + (fy - y1) ** 2)
                r2 = np.sqrt((fx - x2) ** 2 + (fy - y2) ** 2)
                c2 = self.gm * Pc[..., 3] / Pc[..., 0] if self.euler else cs2
                tw = np.sqrt(m1 / (r1 + 1e-12 * dx) ** 3 + m2 / (r2 + 1e-12 * dx) ** 3)
                return val * c2 / tw / np.sqrt(self.gm)

            def tau(Pc, gx, gy, n):
                dvx_dx, dvy_dy = gx[..., 1] / dx, gy[..., 2] / dx
                dvy_dx, dvx_dy = gx[..., 2] / dx, gy[..., 1] / dx
                div = dvx_dx + dvy_dy
                rho = Pc[..., 0]
                txx = rho * n * (2 * dvx_dx - 2.0 / 3 * div) + rho * lam * n * div
                tyy = rho * n * (2 * dvy_dy - 2.0 / 3 * div) + rho * lam * n * div
                txy = rho * n * (dvx_dy + dvy_dx)
                return (txx, txy) if ax == 0 else (txy, tyy)

            nL, nR = nu(PLc), nu(PRc)
            tL, tR = tau(PLc, gLx, gLy, nL), tau(PRc, gRx, gRy, nR)
            tx, ty = 0.5 * (tL[0] + tR[0]), 0.5 * (tL[1] + tR[1])
            F = self.hlle(pl, pr, ax, cs2)
            F[..., 1] -= tx
            F[..., 2] -= ty
            if self.euler:
                F[..., 3] -= tx * 0.5 * (PLc[..., 1] + PRc[..., 1]) + ty * 0.5 * (PLc[..., 2] + PRc[..., 2])
            fluxes.append(F)
        du = -dt * ((fluxes[1] - fluxes[0]) / dx + (fluxes[3] - fluxes[2]) / dx)
        items = self.sources(u, u_init, t, dt, bodies)
        u1 = u + du + sum(items)
        nb, bs = self.mesh.num_blocks, self.mesh.block_size
        ds = np.stack([to_blocks(it, nb, bs).sum(axis=(1, 2)) for it in items], 1) * dx * dx
        return u1, ds

    def sources(self, u, u0, t, dt, bodies):
        ph, x, y = self.ph, self.XC, self.YC
        d = u[..., 0]
        P = self.prim(u)
        vx, vy = P[..., 1], P[..., 2]
        nv = u.shape[-1]
        out = [np.zeros_like(u) for _ in range(7)]
        for k, (m, bx, by, bvx, bvy) in enumerate(bodies):
            t2 = (x - bx) ** 2 + (y - by) ** 2 + self.rs ** 2
            fx, fy = -d * m * (x - bx) / t2 ** 1.5, -d * m * (y - by) / t2 ** 1.5
            out[2 + k][..., 1], out[2 + k][..., 2] = fx * dt, fy * dt
            if self.euler:
                out[2 + k][..., 3] = (fx * vx + fy * vy) * dt
            r2 = (x - bx) ** 2 + (y - by) ** 2
            rate = np.where(r2 < ph.sink_radius ** 2, ph.sink_rate * (1 - r2 / ph.sink_radius ** 2) ** 2, 0.0)
            if ph.sink_model == "impartial":
                out[k] = -rate[..., None] * u * dt
            else:
                if ph.one_body:
                    bvx = bvy = 0.0
                r = np.sqrt(r2)
                rx, ry = (x - bx) / r, (y - by) / r
                dv = (vx - bvx) * rx + (vy - bvy) * ry
                cx, cy = dv * rx + bvx, dv * ry + bvy
                s = [d, d * cx, d * cy]
                if self.euler:
                    e0 = P[..., 3] / d / (self.gm - 1)
                    s.append(d * e0 + 0.5 * d * (cx ** 2 + cy ** 2))
                out[k] = -rate[..., None] * np.stack(s, -1) * dt
        R = self.mesh.domain_radius
        rr = np.sqrt(x ** 2 + y ** 2)
        brate = 0.5 * ph.buffer_rate * (1 + np.tanh((rr - R) / ph.buffer_scale)) * np.sqrt(1.0 / R ** 3)
        out[4] = -(u - u0) * dt * brate[..., None]
        if self.euler:
            h = self.hydro
            e0 = P[..., 3] / d / (self.gm - 1)
            if h.cooling_beta is not None:
                P0 = self.prim(u0)
                et = P0[..., 3] / P0[..., 0] / (self.gm - 1)
                t2 = x ** 2 + y ** 2 + self.rs ** 2
                om = np.sqrt(1.0 / t2 ** 1.5)
                ec = et + (e0 - et) * np.exp(-dt * om / h.cooling_beta)
            else:
                a = h.cooling_coefficient / d ** h.density_index
                ec = e0 * (1 + 3 * a * e0 ** 3 * dt) ** (-1.0 / 3)
            if ph.mach_ceiling is not None:
                ek = 0.5 * (vx ** 2 + vy ** 2)
                ec = np.maximum(ec, 2 * ek / self.gm / (self.gm - 1) / ph.mach_ceiling ** 2)
            Pc = P.copy()
            Pc[..., 3] = d * ec * (self.gm - 1)
            out[5] = self.cons(Pc) - u
        elif ph.fake_mass_rate != 0.0 or ph.fake_mass_threshold != 0.0:
            mask = u[..., 0] < u0[..., 0] * ph.fake_mass_threshold
            out[6] = np.where(mask[..., None], 0.0, u0 * ph.fake_mass_rate * dt)
        return out

    def step(self, u, u_init, ist, t, dt):
        """One full RK step (SSP weights b = 1/2; 3/4, 1/3)."""
        rk = self.ph.rk_substeps()
        u1, d1 = self.stage(u, u_init, t, dt)
        i1 = ist + d1
        if rk == 1:
            return u1, i1, t + dt
        u2, d2 = self.stage(u1, u_init, t + dt, dt)
        i2 = i1 + d2
        if rk == 2:
            return 0.5 * u2 + 0.5 * u, 0.5 * i2 + 0.5 * ist, 0.5 * (t + 2 * dt) + 0.5 * t
        u2, i2, t2 = 0.25 * u2 + 0.75 * u, 0.25 * i2 + 0.75 * ist, 0.25 * (t + 2 * dt) + 0.75 * t
        u3, d3 = self.stage(u2, u_init, t2, dt)
        i3 = i2 + d3
        return (2.0 / 3) * u3 + (1.0 / 3) * u, (2.0 / 3) * i3 + (1.0 / 3) * ist, (2.0 / 3) * (t2 + dt) + t / 3
