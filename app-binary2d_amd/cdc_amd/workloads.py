"""Named workloads: the configurations the parity tests and the benchmark use.

  iso256    isothermal, 4x4 blocks of 64^2 (256^2), RK2, 1 orbit
  iso1024   the same physics at 16x16 blocks of 64^2 (1024^2), 200 steps
  euler2048 gamma-law, 32x32 blocks of 64^2 (2048^2), RK3, alpha viscosity,
            eccentric (e = 0.5) unequal (q = 0.5) binary, beta cooling, pressure
            floor 1e-10 (as the reference's cooling-cbd setup floors pressure), 100 steps
iso2048   the isothermal physics at 32x32 blocks of 64^2 (2048^2): the fine
            run of the learned-closure study (coarse grid 512^2)
euler256  the euler2048 gas and binary at 4x4 blocks of 64^2 (256^2) with
            constant nu = 1e-3 in place of alpha viscosity: the gamma-law
            configuration that runs a whole orbit (1063 RK3 steps); alpha
            viscosity outruns its explicit limit in the hot cavity gas
ensemble64 (ensemble_members) 64 isothermal RK2 runs of 256^2 to 1024^2 with
            drawn binary and disk parameters, half an orbit each
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
    "iso2048": dict(mesh=Mesh(32, 64, 8.0), hydro=Isothermal(mach_number=10.0), physics=base_physics(),
                    seed=5, steps=None, orbits=None),
    "euler2048": dict(mesh=Mesh(32, 64, 8.0),
                      hydro=Euler(gamma_law_index=5.0 / 3.0, cooling_coefficient=0.0, cooling_beta=10.0,
                                  pressure_floor=1e-10),
                      physics=base_physics(binary_eccentricity=0.5, binary_mass_ratio=0.5, rk_order="RK3",
                                           viscosity={"alpha": 0.1}),
                      seed=3, steps=100, orbits=None),
    "euler256": dict(mesh=Mesh(4, 64, 8.0),
                     hydro=Euler(gamma_law_index=5.0 / 3.0, cooling_coefficient=0.0, cooling_beta=10.0,
                                 pressure_floor=1e-10),
                     physics=base_physics(binary_eccentricity=0.5, binary_mass_ratio=0.5, rk_order="RK3",
                                          viscosity={"constant_nu": 1e-3}),
                     seed=3, steps=None, orbits=1.0),
}


def ensemble_members(seed=4, counts=((24, 4), (24, 8), (16, 16)), block_size=64):
    """The 64-run parameter study: 24 runs at 256^2, 24 at 512^2, 16 at 1024^2
    (blocks of 64^2), isothermal RK2, each with binary eccentricity
    e ~ U[0, 0.8], mass ratio q ~ log-U[0.1, 1], Mach number ~ U[10, 40] and
    kinematic viscosity nu ~ log-U[1e-4, 1e-2] drawn from one seeded
    generator; every run has its own initial-state seed."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for count, nb in counts:
        for _ in range(count):
            e = float(rng.uniform(0.0, 0.8))
            q = float(10.0 ** rng.uniform(-1.0, 0.0))
            mach = float(rng.uniform(10.0, 40.0))
            nu = float(10.0 ** rng.uniform(-4.0, -2.0))
            out.append(dict(mesh=Mesh(nb, block_size, 8.0), hydro=Isothermal(mach_number=mach),
                            physics=base_physics(binary_eccentricity=e, binary_mass_ratio=q,
                                                 viscosity={"constant_nu": nu}),
                            seed=1000 * seed + len(out), steps=None, orbits=0.5))
    return out


def training_runs(count, seed=6, blocks=32, block_size=64):
    """Fine isothermal runs for training the learned closure: the iso2048
    mesh (32 x 32 blocks of 64^2) with binary and gas parameters drawn from
    the 64-run study's ranges (e ~ U[0, 0.8], q ~ log-U[0.1, 1], Mach ~
    U[10, 40], nu ~ log-U[1e-4, 1e-2]) by a seeded generator, each with its
    own initial-state seed."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        e = float(rng.uniform(0.0, 0.8))
        q = float(10.0 ** rng.uniform(-1.0, 0.0))
        mach = float(rng.uniform(10.0, 40.0))
        nu = float(10.0 ** rng.uniform(-4.0, -2.0))
        out.append(dict(mesh=Mesh(blocks, block_size, 8.0), hydro=Isothermal(mach_number=mach),
                        physics=base_physics(binary_eccentricity=e, binary_mass_ratio=q,
                                             viscosity={"constant_nu": nu}),
                        seed=1000 * seed + k, steps=None, orbits=None))
    return out


def get(name, **over):
    """Deep copy of a named workload; keyword overrides replace top-level keys."""
    w = copy.deepcopy(WORKLOADS[name])
    w.update(over)
    return w


def initial_state(w):
    from .solver import seeded_cavity
    return seeded_cavity(w["mesh"], w["hydro"], softening_length=w["physics"].softening(),
                         mach_number=10.0, seed=w["seed"], taper_radius=TAPER_RADIUS, taper_width=TAPER_WIDTH)
