"""Named workloads: the configurations the parity tests and the benchmark use.

  iso256    isothermal, 4x4 blocks of 64^2 (256^2), RK2, 1 orbit
  iso1024   the same physics at 16x16 blocks of 64^2 (1024^2), 200 steps
  euler2048 gamma-law, 32x32 blocks of 64^2 (2048^2), RK3, alpha viscosity,
            eccentric (e = 0.5) unequal (q = 0.5) binary, beta cooling, pressure
            floor 1e-10 (as the reference's cooling-cbd setup floors pressure), 100 steps
Domain [-8a, 8a]^2, a = M = G = 1, softening = sink radius = 0.05, sink rate
10, Mach 10, PLM theta 1.5, CFL 0.4, buffer (rate 1000, scale 0.3) relaxing to
the initial state near the domain edge; initial state: the seeded cavity
model (cavity radius 2.5a, floor 1e-4, noise 1e-3, Keplerian flow tapered to
rest outside r = 7a).
"""
import copy
import math

from .config import Euler, Isothermal, Mesh, Physics

ORBITAL_PERIOD = 2.0 * math.pi
# The Keplerian flow is tapered to rest outside r = 7a (inside the buffer
# zone): a rotating disk is not periodic, and the shear where the periodic
# mesh wraps heats the gas until alpha viscosity outruns its explicit limit.
TAPER_RADIUS = 7.0
TAPER_WIDTH = 0.25


def base_physics(**kw):
    p = Physics(buffer_rate=1000.0, buffer_scale=0.3, binary_eccentricity=0.0, binary_mass_ratio=1.0,
                cfl=0.4, plm=1.5, rk_order="RK2", sink_radius=0.05, sink_rate=10.0,
                softening_length=0.05, viscosity={"constant_nu": 1e-3})
    for k, v in kw.items():
        setattr(p, k, v)
    return p


WORKLOADS = {
    "iso256": dict(mesh=Mesh(4, 64, 8.0), hydro=Isothermal(mach_number=10.0), physics=base_physics(),
                   seed=1, steps=None, orbits=1.0),
    "iso1024": dict(mesh=Mesh(16, 64, 8.0), hydro=Isothermal(mach_number=10.0), physics=base_physics(),
                    seed=1, steps=200, orbits=None),
    "euler2048": dict(mesh=Mesh(32, 64, 8.0),
                      hydro=Euler(gamma_law_index=5.0 / 3.0, cooling_coefficient=0.0, cooling_beta=10.0,
                                  pressure_floor=1e-10),
                      physics=base_physics(binary_eccentricity=0.5, binary_mass_ratio=0.5, rk_order="RK3",
                                           viscosity={"alpha": 0.1}),
                      seed=3, steps=100, orbits=None),
}


def get(name, **over):
    """Deep copy of a named workload; keyword overrides replace top-level keys."""
    w = copy.deepcopy(WORKLOADS[name])
    w.update(over)
    return w


def initial_state(w):
    from .solver import seeded_cavity
    return seeded_cavity(w["mesh"], w["hydro"], softening_length=w["physics"].softening(),
                         mach_number=10.0, seed=w["seed"], taper_radius=TAPER_RADIUS, taper_width=TAPER_WIDTH)
