"""The reference's convergence harness, restated: the residual test
(tools/residual_test, setups/residual_test.yaml) runs the Euler scheme at two
resolutions on the smooth ResidualTestModel (src/disks.rs:208-229) and
measures how well each discrete solution satisfies the PDE, with an
independent finite-difference form of the equations
(tools/indep_res_analysis.py): the L2 norm, outside the buffer zone, of

    R = dU/dt + div F_inviscid - div F_viscous - S

with dU/dt the centred difference of consecutive steps, div the centred
difference on the periodic mesh, and S the gravity, buffer, cooling and sink
sources at the middle step.  The residual falls as dx^p with p about 1.5;
the harness accepts 1.25 .. 2 (tools/residual_test:52-54).

The sources follow the solver's own forms (src/physics.rs): the sink kernel
rate (1 - r^2/s^2)^2 inside the sink radius (src/physics.rs:388-397), the
impartial sink, T^4 cooling d(Sigma e)/dt = -a e^4 Sigma with
a = coefficient / Sigma^2 (density_index 2, src/physics.rs:721-728),
Plummer gravity with the softening length, and the buffer relaxation toward
the initial state (src/physics.rs:443-446).
"""
import numpy as np

from cdc_amd import Euler, Mesh, Physics

GAMMA = 1.66666666666666
T_FINAL = 0.05 * 2.0 * np.pi   # control.num_orbits = 0.05 of the short test


def configuration(block_size):
    """setups/residual_test.yaml with mesh.block_size = block_size."""
    mesh = Mesh(4, block_size, 6.0)
    hydro = Euler(gamma_law_index=GAMMA, cooling_coefficient=0.025, density_index=2,
                  cooling_slow_start=False)
    physics = Physics(buffer_rate=1000.0, buffer_scale=0.3, binary_eccentricity=0.0,
                      binary_mass_ratio=0.5, cfl=0.1, plm=1.8, rk_order="RK2", sink_radius=1.5,
                      sink_rate=1.0, lambda_multiplier=1.66666666666666,
                      viscosity={"alpha": 0.001}, softening_length=1.0)
    return mesh, hydro, physics


def model_primitive(x, y):
    """ResidualTestModel::primitive_at (src/disks.rs:209-226): Sigma, vx, vy, p."""
    r = np.sqrt(x * x + y * y)
    phi = np.arctan2(x, y)
    r1 = np.sqrt((x - 1.0) ** 2 + (y - 1.0) ** 2)
    r2 = np.sqrt((x + 1.0) ** 2 + (y + 1.0) ** 2)
    sd = 1.0 + np.exp(-r1 ** 2)
    pg = 1.0 + np.exp(-r2 ** 2)
    env = np.exp(-5.0 / r - r ** 2 / 3.0)
    vp = r ** -0.5 * env
    vr = np.sin(phi - 3.14159 / 4.0) * env
    vx = vp * (-y / r) + vr * (x / r)
    vy = vp * (x / r) + vr * (y / r)
    return sd, vx, vy, pg


def initial_conserved(mesh):
    """[block][i][j][4] conserved state of the model at the cell centres."""
    out = np.empty((mesh.num_blocks ** 2, mesh.block_size, mesh.block_size, 4))
    for b, idx in enumerate(mesh.block_indexes()):
        xc, yc = mesh.cell_centers(idx)
        sd, vx, vy, pg = model_primitive(xc, yc)
        out[b, ..., 0] = sd
        out[b, ..., 1] = sd * vx
        out[b, ..., 2] = sd * vy
        out[b, ..., 3] = pg / (GAMMA - 1.0) + 0.5 * sd * (vx * vx + vy * vy)
    return out


def to_global(mesh, u):
    """[block][i][j][v] -> [N][N][v] (block (bi, bj) at rows bi*bs, cols bj*bs)."""
    nb, bs = mesh.num_blocks, mesh.block_size
    return u.reshape(nb, nb, bs, bs, -1).transpose(0, 2, 1, 3, 4).reshape(nb * bs, nb * bs, -1)


def global_centres(mesh):
    n = mesh.num_blocks * mesh.block_size
    x = np.empty(n)
    for bi in range(mesh.num_blocks):
        xc, _ = mesh.cell_centers((bi, 0))
        x[bi * mesh.block_size:(bi + 1) * mesh.block_size] = xc[:, 0]
    return np.meshgrid(x, x, indexing="ij")


class Residual:
    """L2 residuals of the four conservation laws at the middle of three
    consecutive states (x is axis 0, y axis 1, periodic).  Called with the
    states, it returns one [4] norm vector per entry of `drops`: None for the
    full residual, "gravity" for the residual without the gravity source
    (to show that the measure notices a missing term)."""

    def __init__(self, mesh, hydro, physics, u_init, drops=(None,)):
        self.mesh, self.hydro, self.physics, self.drops = mesh, hydro, physics, drops
        self.dx = mesh.cell_spacing()
        self.x, self.y = global_centres(mesh)
        r = np.sqrt(self.x ** 2 + self.y ** 2)
        R = mesh.domain_radius
        self.inside = r <= R - physics.buffer_scale
        self.ncell = r.size
        self.u_init = [np.ascontiguousarray(p) for p in np.moveaxis(to_global(mesh, u_init), -1, 0)]
        om = np.sqrt(1.0 / R ** 3)
        self.buffer = 0.5 * physics.buffer_rate * (1.0 + np.tanh((r - R) / physics.buffer_scale)) * om

    def ddx(self, f):
        return (np.roll(f, -1, axis=0) - np.roll(f, 1, axis=0)) / (2.0 * self.dx)

    def ddy(self, f):
        return (np.roll(f, -1, axis=1) - np.roll(f, 1, axis=1)) / (2.0 * self.dx)

    def __call__(self, u_prev, t_prev, u, t, u_next, t_next, bodies):
        """u_*: [N][N][4]; bodies: [[m, x, y, vx, vy], [m, x, y, vx, vy]] at t."""
        g = GAMMA
        ph = self.physics
        rho, sx, sy, E = (np.ascontiguousarray(u[..., v]) for v in range(4))
        vx, vy = sx / rho, sy / rho
        p = (g - 1.0) * (E - 0.5 * rho * (vx * vx + vy * vy))
        eps = p / ((g - 1.0) * rho)
        idt = 1.0 / (t_next - t_prev)

        (m1, x1, y1, _, _), (m2, x2, y2, _, _) = bodies
        d1x, d1y, d2x, d2y = self.x - x1, self.y - y1, self.x - x2, self.y - y2
        r1s, r2s = d1x * d1x + d1y * d1y, d2x * d2x + d2y * d2y

        # alpha viscosity: nu = alpha c_s^2 / sqrt(gamma) / sqrt(m1/r1^3 + m2/r2^3)
        alpha = ph.viscosity_model()[1]
        nu = alpha * (g * p / rho) / np.sqrt(g) / np.sqrt(m1 / (r1s * np.sqrt(r1s)) + m2 / (r2s * np.sqrt(r2s)))
        lam = ph.lambda_multiplier * nu
        gxx, gyy = self.ddx(vx), self.ddy(vy)
        gxy, gyx = self.ddx(vy), self.ddy(vx)
        div = gxx + gyy
        txx = rho * nu * (2.0 * gxx - (2.0 / 3.0) * div) + rho * lam * div
        tyy = rho * nu * (2.0 * gyy - (2.0 / 3.0) * div) + rho * lam * div
        txy = rho * nu * (gyx + gxy)
        fx = (sx, sx * vx + p - txx, sy * vx - txy, (E + p) * vx - (vx * txx + vy * txy))
        fy = (sy, sx * vy - txy, sy * vy + p - tyy, (E + p) * vy - (vx * txy + vy * tyy))

        rs2 = ph.softening() ** 2
        q1, q2 = r1s + rs2, r2s + rs2
        f1 = -rho * m1 / (q1 * np.sqrt(q1))
        f2 = -rho * m2 / (q2 * np.sqrt(q2))
        gx, gy = f1 * d1x + f2 * d2x, f1 * d1y + f2 * d2y
        grav = (None, gx, gy, vx * gx + vy * gy)
        s2 = ph.sink_radius ** 2
        w = np.zeros_like(rho)
        for rks in (r1s, r2s):
            a = np.maximum(1.0 - rks / s2, 0.0)   # (1 - r^2/s^2)^2 inside, 0 outside
            w += ph.sink_rate * (a * a)
        e2 = eps * eps
        out = [np.empty(4) for _ in self.drops]
        for v, uv in enumerate((rho, sx, sy, E)):
            res = (u_next[..., v] - u_prev[..., v]) * idt + self.ddx(fx[v]) + self.ddy(fy[v])
            res += self.buffer * (uv - self.u_init[v]) + w * uv
            if v == 3:
                res += self.hydro.cooling_coefficient * (e2 * e2) / rho
            for k, d in enumerate(self.drops):
                r = res if (d == "gravity" or grav[v] is None) else res - grav[v]
                out[k][v] = np.sqrt(np.sum(r[self.inside] ** 2) / self.ncell)
        return out if len(self.drops) > 1 else out[0]


def residual_series(mesh, hydro, physics, u_init, advance, state, bodies_at, drops=(None,), stride=1):
    """Residual norms [k][4] and their times from a run to T_FINAL, one
    series per entry of drops (see Residual), at every stride-th step.
    advance(): one RK step (fold 1); state() -> (t, [block][i][j][4])."""
    res = Residual(mesh, hydro, physics, u_init, drops)
    t0, u0 = state()
    hist = [(t0, to_global(mesh, u0))]
    times, norms = [], [[] for _ in drops]
    step = 0
    while hist[-1][0] < T_FINAL:
        advance()
        step += 1
        t, u = state()
        hist.append((t, to_global(mesh, u)))
        if len(hist) == 3:
            (tp, up), (tm, um), (tn, un) = hist
            if (step - 2) % stride == 0:
                r = res(up, tp, um, tm, un, tn, bodies_at(tm))
                for out, x in zip(norms, r if len(drops) > 1 else [r]):
                    out.append(x)
                times.append(tm)
            hist.pop(0)
    times = np.array(times)
    series = [(times, np.array(n)) for n in norms]
    return series[0] if len(drops) == 1 else series


def convergence_orders(low, high):
    """Mean over the low-resolution times of log2(res_low / res_high), the
    high-resolution series interpolated to those times; per conservation
    law (mass, x-momentum, y-momentum, energy)."""
    (tl, rl), (th, rh) = low, high
    keep = (tl >= th[0]) & (tl <= th[-1])
    out = []
    for v in range(4):
        hi = np.interp(tl[keep], th, rh[:, v])
        out.append(float(np.mean(np.log2(rl[keep, v] / hi))))
    return out
