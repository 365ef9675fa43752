"""Shared test helpers: small configurations and comparison norms."""
import copy

import numpy as np

from cdc_amd import Euler, Isothermal, Mesh, workloads
from cdc_amd.solver import seeded_cavity
from oracle import oracle as O


def small(kind="iso", nb=2, bs=16, R=8.0, **phys):
    """A small configuration of the seeded-cavity disk on nb x nb blocks of bs^2."""
    mesh = Mesh(nb, bs, R)
    if kind == "iso":
        hydro = Isothermal(mach_number=10.0)
    elif kind == "euler":
        hydro = Euler(gamma_law_index=5.0 / 3.0, cooling_coefficient=1e-8)
    elif kind == "euler_beta":
        hydro = Euler(gamma_law_index=5.0 / 3.0, cooling_coefficient=0.0, cooling_beta=10.0)
    else:
        raise ValueError(kind)
    p = workloads.base_physics(sink_radius=0.3, softening_length=0.3, sink_rate=1.0, buffer_rate=10.0)
    if kind != "iso":
        p.cfl = 0.2     # the cold gamma-law gas near the softened bodies needs it at dx = R/8
    for k, v in phys.items():
        setattr(p, k, v)
    return mesh, hydro, p


def initial(mesh, hydro, physics, seed=7, noise=1e-3, cavity=2.5, scale=1e-3, mach=3.0):
    """Seeded cavity state with every conserved field scaled by `scale`.  The
    hydro is invariant under that scaling; it keeps the per-block momentum
    exchange with the binary small enough that the orbital-element
    perturbation stays bound on these coarse test meshes."""
    return scale * seeded_cavity(mesh, hydro, cavity_radius=cavity, noise_amplitude=noise,
                                 softening_length=physics.softening(), seed=seed, mach_number=mach)


def oracle_for(mesh, hydro, physics, u0):
    m, h, p = mesh.to_c(), hydro.to_c(), physics.to_c()
    o = O.Oracle(m, h, p, u0, hydro.nv)
    o._keep = (m, h, p)
    return o


def field_err(a, b):
    """max over variables of max|a_v - b_v| / max|b_v| (field-scaled relative error)."""
    a = np.asarray(a)
    b = np.asarray(b)
    nv = b.shape[-1]
    errs = []
    for v in range(nv):
        s = np.max(np.abs(b[..., v]))
        d = np.max(np.abs(a[..., v] - b[..., v]))
        errs.append(d / s if s > 0 else d)
    return max(errs)


def items_err(a, b):
    """ist [blk][7][nv] comparison: per (item, component) scaled by the largest
    |value| of that item/component over blocks, but never by less than 1e-6 of
    the component's largest item (some items are pure rounding noise, such as
    the momentum part of Euler cooling, uc - u with uc = Sigma * (S / Sigma))."""
    a = np.asarray(a)
    b = np.asarray(b)
    s = np.max(np.abs(b), axis=0, keepdims=True)
    s = np.maximum(s, 1e-6 * np.max(s, axis=1, keepdims=True))
    d = np.abs(a - b)
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.where(s > 0, d / np.where(s > 0, s, 1.0), d)
    return float(np.max(r))


def deep(x):
    return copy.deepcopy(x)
