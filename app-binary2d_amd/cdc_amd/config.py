"""Configuration types mirroring the reference's serde structs.

Mesh (src/mesh.rs:19-37), Isothermal (src/physics.rs:258-265), Euler
(src/physics.rs:273-303), Physics (src/physics.rs:123-236) and the YAML
Configuration (src/app.rs:169-177).  Field names, defaults, Option
semantics and the rejection of unknown fields (serde deny_unknown_fields)
follow the reference, so the reference's setups/*.yaml load unchanged.
Each type converts to its C-ABI struct with `to_c()`.
"""
from dataclasses import dataclass, field, fields
from typing import Optional

import numpy as np
import yaml

from . import abi


def _from_dict(cls, d, what):
    if d is None:
        d = {}
    names = {f.name for f in fields(cls)}
    alias = getattr(cls, "_aliases", {})
    kw = {}
    for k, v in d.items():
        k2 = alias.get(k, k)
        if k2 not in names:
            raise ValueError(f"unknown field `{k}` in {what}")
        kw[k2] = v
    return cls(**kw)


@dataclass
class Mesh:
    """Mesh, src/mesh.rs:19-111."""
    num_blocks: int
    block_size: int
    domain_radius: float

    def block_length(self):
        return 2.0 * self.domain_radius / float(self.num_blocks)

    def cell_spacing(self):
        return self.block_length() / float(self.block_size)

    def total_zones(self):
        return self.num_blocks * self.num_blocks * self.block_size * self.block_size

    def block_indexes(self):
        return [(i, j) for i in range(self.num_blocks) for j in range(self.num_blocks)]

    def block_vertices(self, index):
        """linspace(start, start + L, bs + 1) per axis (src/mesh.rs:52-57)."""
        L = self.block_length()
        out = []
        for b in index:
            a = -self.domain_radius + float(b) * L
            step = ((a + L) - a) / float(self.block_size)
            out.append(a + step * np.arange(self.block_size + 1, dtype=np.float64))
        return tuple(out)

    def cell_centers(self, index):
        xv, yv = self.block_vertices(index)
        xc = (xv[:-1] + xv[1:]) / 2.0
        yc = (yv[:-1] + yv[1:]) / 2.0
        return np.meshgrid(xc, yc, indexing="ij")

    def to_c(self):
        return abi.CMesh(int(self.num_blocks), int(self.block_size), float(self.domain_radius))


@dataclass
class Isothermal:
    """Isothermal, src/physics.rs:258-265."""
    mach_number: float
    density_floor: Optional[float] = None
    nv = 3

    def to_c(self):
        h = abi.CHydro()
        h.kind = abi.CDC_ISOTHERMAL
        h.mach_number = float(self.mach_number)
        h.density_index = 2
        h.has_density_floor = int(self.density_floor is not None)
        h.density_floor = float(self.density_floor or 0.0)
        return h


@dataclass
class Euler:
    """Euler, src/physics.rs:273-307.  `cooling_beta`, when set, selects the
    beta-cooling law (t_cool = beta / Omega), which has no reference
    counterpart; otherwise the reference's T^4 law applies."""
    gamma_law_index: float
    cooling_coefficient: float
    pressure_floor: Optional[float] = None
    density_floor: Optional[float] = None
    density_index: int = 2
    cooling_slow_start: bool = False
    cooling_beta: Optional[float] = None
    nv = 4

    def to_c(self):
        h = abi.CHydro()
        h.kind = abi.CDC_EULER
        h.gamma_law_index = float(self.gamma_law_index)
        h.cooling_coefficient = float(self.cooling_coefficient)
        h.density_index = int(self.density_index)
        h.cooling_slow_start = int(bool(self.cooling_slow_start))
        h.has_density_floor = int(self.density_floor is not None)
        h.density_floor = float(self.density_floor or 0.0)
        h.has_pressure_floor = int(self.pressure_floor is not None)
        h.pressure_floor = float(self.pressure_floor or 0.0)
        h.cooling_model = abi.CDC_COOLING_BETA if self.cooling_beta is not None else abi.CDC_COOLING_T4
        h.cooling_beta = float(self.cooling_beta or 0.0)
        return h


def hydro_from_dict(d):
    """AnyHydro, externally tagged (src/app.rs:28-33)."""
    if not isinstance(d, dict) or len(d) != 1:
        raise ValueError("hydro must have exactly one of `isothermal`, `euler`")
    (k, v), = d.items()
    if k == "isothermal":
        return _from_dict(Isothermal, v, "hydro.isothermal")
    if k == "euler":
        return _from_dict(Euler, v, "hydro.euler")
    raise ValueError(f"unknown hydro `{k}`")


RK_ORDERS = {"RK1": 1, "RK2": 2, "RK3": 3}


@dataclass
class Physics:
    """Physics, src/physics.rs:123-236."""
    buffer_rate: float
    buffer_scale: float
    binary_eccentricity: float
    binary_mass_ratio: float
    cfl: float
    plm: float
    rk_order: str
    sink_radius: float
    sink_rate: float
    lambda_multiplier: float = 0.0
    viscosity: dict = field(default_factory=lambda: {"constant_nu": 1e-3})
    sink_model: str = "impartial"
    plm_density_trigger: Optional[float] = None
    fake_mass_threshold: float = 0.0
    fake_mass_rate: float = 0.0
    one_body: bool = False
    mach_ceiling: Optional[float] = None
    softening_length: Optional[float] = None
    dt_min: Optional[float] = None
    safe_mode_duration: float = 0.0
    safe_cfl: Optional[float] = None
    safe_plm: Optional[float] = None
    safe_mach_ceiling: Optional[float] = None
    safe_rk_order: Optional[str] = None
    lambda_: Optional[float] = None
    nu: Optional[float] = None
    alpha: Optional[float] = None
    _aliases = {"lambda": "lambda_"}

    def viscosity_model(self):
        """Physics::viscosity_model, src/physics.rs:222-230 (deprecated fields win)."""
        if self.nu is not None:
            return ("constant_nu", float(self.nu))
        if self.alpha is not None:
            return ("alpha", float(self.alpha))
        (k, v), = self.viscosity.items()
        if k not in ("constant_nu", "alpha"):
            raise ValueError(f"unknown viscosity model `{k}`")
        return (k, float(v))

    def softening(self):
        return self.sink_radius if self.softening_length is None else self.softening_length

    def rk_substeps(self):
        return RK_ORDERS[self.rk_order]

    def safe(self):
        """The safe-mode variant of src/main.rs:131-147."""
        import copy
        p = copy.deepcopy(self)
        if self.safe_cfl is not None:
            p.cfl = self.safe_cfl
        if self.safe_plm is not None:
            p.plm = self.safe_plm
        if self.safe_mach_ceiling is not None:
            p.mach_ceiling = self.safe_mach_ceiling
        if self.safe_rk_order is not None:
            p.rk_order = self.safe_rk_order
        return p

    def to_c(self):
        c = abi.CPhysics()
        c.buffer_rate = float(self.buffer_rate)
        c.buffer_scale = float(self.buffer_scale)
        c.binary_eccentricity = float(self.binary_eccentricity)
        c.binary_mass_ratio = float(self.binary_mass_ratio)
        c.lambda_multiplier = float(self.lambda_multiplier)
        kind, val = self.viscosity_model()
        c.viscosity_model = abi.CDC_VISC_ALPHA if kind == "alpha" else abi.CDC_VISC_CONSTANT_NU
        c.viscosity = val
        if self.sink_model not in ("impartial", "torque_free"):
            raise ValueError(f"unknown sink model `{self.sink_model}`")
        c.sink_model = abi.CDC_SINK_TORQUE_FREE if self.sink_model == "torque_free" else abi.CDC_SINK_IMPARTIAL
        c.cfl = float(self.cfl)
        c.plm = float(self.plm)
        c.has_plm_density_trigger = int(self.plm_density_trigger is not None)
        c.plm_density_trigger = float(self.plm_density_trigger or 0.0)
        c.one_body = int(bool(self.one_body))
        c.fake_mass_threshold = float(self.fake_mass_threshold)
        c.fake_mass_rate = float(self.fake_mass_rate)
        c.has_mach_ceiling = int(self.mach_ceiling is not None)
        c.mach_ceiling = float(self.mach_ceiling or 0.0)
        c.rk_order = RK_ORDERS[self.rk_order]
        c.sink_radius = float(self.sink_radius)
        c.sink_rate = float(self.sink_rate)
        c.has_softening_length = int(self.softening_length is not None)
        c.softening_length = float(self.softening_length or 0.0)
        c.has_dt_min = int(self.dt_min is not None)
        c.dt_min = float(self.dt_min or 0.0)
        return c


@dataclass
class Configuration:
    """Configuration, src/app.rs:169-177 (model and control kept as dicts)."""
    hydro: object
    model: dict
    mesh: Mesh
    physics: Physics
    control: dict

    @classmethod
    def from_dict(cls, d):
        extra = set(d) - {"hydro", "model", "mesh", "physics", "control"}
        if extra:
            raise ValueError(f"unknown configuration fields {sorted(extra)}")
        return cls(hydro=hydro_from_dict(d["hydro"]), model=d.get("model", {}),
                   mesh=_from_dict(Mesh, d["mesh"], "mesh"),
                   physics=_from_dict(Physics, d["physics"], "physics"),
                   control=d.get("control", {}))

    @classmethod
    def from_yaml(cls, path):
        with open(path) as f:
            return cls.from_dict(yaml.safe_load(f))
