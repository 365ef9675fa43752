"""Checkpoints in the reference's App layout (src/app.rs:185-192), as CBOR.

The reference writes `App {state, tasks, time_series, config, version}` with
serde + ciborium (src/io.rs:7-18).  This module writes and reads the same
tree with its own CBOR codec, following serde's data model:
  structs        -> maps keyed by field name
  enums          -> externally tagged maps {"variant": value} (snake_case)
  State          -> {"time", "iteration": [numer, denom], "solution": {[i, j]: BlockState}}
                    (Rational64 as a (numer, denom) tuple; the solution map
                    keyed by the BlockIndex tuple)
  BlockState     -> {"conserved": {"v": 1, "dim": [bs, bs], "data": [[u0..u_nv-1] x bs^2]},
                     "integrated_source_terms": ItemizedChange<Conserved>,
                     "orbital_elements_change": ItemizedChange<OrbitalElements>}
                    (ndarray's serde form; Conserved and OrbitalElements as
                    tuples)
  ItemizedChange -> {"sink1", "sink2", "grav1", "grav2", "buffer", "cooling", "fake_mass"}
  Tasks          -> {"write_checkpoint": {"count", "next_time"}, ...} (src/tasks.rs:10-50)
  TimeSeries     -> {"isothermal" | "euler": [{"time", "integrated_source_terms",
                                               "orbital_elements_change"}, ...]}
                    (plus "torque": [tau_1, tau_2] per sample when recorded, an
                    extension field that serde skips)
The reference reads `config` through structs that reject unknown fields
(Configuration, Physics, Isothermal, Euler, and the AnyHydro / AnyModel
enums: src/app.rs:26-48, src/app.rs:169-177, src/physics.rs:121-123,
src/physics.rs:255-273), so `config` carries exactly the reference's serde
fields (REF_FIELDS below) and `model` is always a valid AnyModel variant.
What has no place there goes under the App-level key "cdc_amd" (App itself
accepts unknown keys, src/app.rs:182-192): the beta-cooling parameter, an
initial model that is not one of the reference's (the seeded cavity; the
config then names `infinite_disk` with surface density 1, the disk the
seeded cavity has outside its cavity, which a reference restart uses as its
buffer-zone target), and the per-block torque impulses, so that a restart
carries the torque totals on.  The layout is restated from those
conventions; no reference checkpoint file exists to pin it, so it is
covered by round trips and by a check of every config key against
REF_FIELDS.
"""
import dataclasses

import numpy as np

from . import cbor

ITEMS = ("sink1", "sink2", "grav1", "grav2", "buffer", "cooling", "fake_mass")

# The reference's serde field names of each configuration struct.
REF_FIELDS = {
    # Mesh, src/mesh.rs:19-37
    "mesh": ("num_blocks", "block_size", "domain_radius"),
    # Isothermal, src/physics.rs:255-265
    "isothermal": ("mach_number", "density_floor"),
    # Euler, src/physics.rs:270-303
    "euler": ("gamma_law_index", "cooling_coefficient", "pressure_floor", "density_floor", "density_index",
              "cooling_slow_start"),
    # Physics, src/physics.rs:121-217
    "physics": ("buffer_rate", "buffer_scale", "binary_eccentricity", "binary_mass_ratio", "lambda_multiplier",
                "viscosity", "sink_model", "cfl", "plm", "plm_density_trigger", "fake_mass_threshold",
                "fake_mass_rate", "one_body", "mach_ceiling", "rk_order", "sink_radius", "sink_rate",
                "softening_length", "dt_min", "safe_mode_duration", "safe_cfl", "safe_plm", "safe_mach_ceiling",
                "safe_rk_order", "lambda", "nu", "alpha"),
    # Control, src/app.rs:94-131
    "control": ("num_orbits", "checkpoint_interval", "time_series_interval", "fold", "num_threads",
                "output_directory", "fallback_stack_size"),
}
# AnyModel variants (src/app.rs:41-48) and each model's fields (src/disks.rs:11-93)
REF_MODELS = {
    "infinite_disk": ("mach_number", "surface_density"),
    "infinite_alpha_disk": ("mach_number", "surface_density"),
    "finite_disk": ("disk_mass", "disk_width", "disk_radius", "softening_length", "mach_number"),
    "residual_test": (),
}
EXT = "cdc_amd"
# Control defaults of a run that names none (the reference requires the fields)
CONTROL_DEFAULTS = {"num_orbits": 1.0, "checkpoint_interval": 1.0, "time_series_interval": None, "fold": 1,
                    "num_threads": None, "output_directory": "data", "fallback_stack_size": 0}
TASKS = ("write_checkpoint", "iteration_message", "record_time_series", "report_progress",
         "simulation_startup")
VERSION = "cdc-amd checkpoint 1"


def _items(a):
    """[7][k] -> ItemizedChange map of k-tuples."""
    return {name: [float(v) for v in a[i]] for i, name in enumerate(ITEMS)}


def _unitems(d, k):
    return np.array([[float(v) for v in d[name]] for name in ITEMS]).reshape(7, k)


def _plain(x):
    """dataclass / dict tree -> serde-like plain tree (None stays null)."""
    if dataclasses.is_dataclass(x):
        return {f.name: _plain(getattr(x, f.name)) for f in dataclasses.fields(x)}
    if isinstance(x, dict):
        return {k: _plain(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [_plain(v) for v in x]
    return x


def hydro_tag(hydro):
    return "euler" if getattr(hydro, "nv", 3) == 4 else "isothermal"


def new_tasks():
    return {name: {"count": 0, "next_time": 0.0} for name in TASKS}


def _ref_struct(x, which):
    """The reference's serde fields of a config object (dataclass or dict),
    in REF_FIELDS order; this package's field `lambda_` is the reference's
    `lambda`."""
    d = _plain(x) if not isinstance(x, dict) else dict(x)
    if "lambda_" in d:
        d["lambda"] = d.pop("lambda_")
    if which == "control":
        d = {**CONTROL_DEFAULTS, **{k: v for k, v in d.items() if v is not None or k in CONTROL_DEFAULTS}}
    return {k: d[k] for k in REF_FIELDS[which] if k in d}


def _ref_model(model, hydro):
    """(config model, extension model): a reference AnyModel passes through;
    anything else is kept under the extension key and the config names the
    infinite disk of surface density 1 (with a Mach number only where the
    hydro has none: the reference accepts exactly one)."""
    if isinstance(model, dict) and len(model) == 1:
        (k, v), = model.items()
        if k in REF_MODELS and set(v or {}) <= set(REF_MODELS[k]):
            return {k: dict(v or {})}, None
    mach = None if getattr(hydro, "nv", 3) == 3 else float(getattr(hydro, "mach_number", 10.0))
    if isinstance(model, dict) and len(model) == 1:
        (_, v), = model.items()
        if isinstance(v, dict) and getattr(hydro, "nv", 3) == 4 and "mach_number" in v:
            mach = float(v["mach_number"])
    return {"infinite_disk": {"mach_number": mach, "surface_density": 1.0}}, (model or None)


def pack(state, mesh, hydro, physics, model=None, control=None, tasks=None, time_series=None, torque=None):
    """The App tree for a state dict (time, iteration, conserved [block][i][j][v],
    ist [block][7][nv], oec [block][7][4]); `torque` [block][2] are the
    per-block torque impulses (Solver.torque())."""
    nb, bs = mesh.num_blocks, mesh.block_size
    u, ist, oec = state["conserved"], state["ist"], state["oec"]
    nv = u.shape[-1]
    solution = {}
    for b in range(nb * nb):
        solution[(b // nb, b % nb)] = {
            "conserved": {"v": 1, "dim": [bs, bs], "data": cbor.Rows(u[b].reshape(bs * bs, nv))},
            "integrated_source_terms": _items(ist[b]),
            "orbital_elements_change": _items(oec[b]),
        }
    tag = hydro_tag(hydro)
    num, den = state["iteration"]
    series = []
    for s in time_series or []:
        e = {"time": float(s["time"]), "integrated_source_terms": _items(s["ist"]),
             "orbital_elements_change": _items(s["oec"])}
        if "torque" in s:      # extension field; serde ignores unknown fields here
            e["torque"] = [float(v) for v in s["torque"]]
        series.append(e)
    cmodel, emodel = _ref_model(model, hydro)
    ext = {}
    if emodel is not None:
        ext["model"] = emodel
    if getattr(hydro, "cooling_beta", None) is not None:
        ext["cooling_beta"] = float(hydro.cooling_beta)
    if torque is not None:
        ext["torque"] = cbor.Rows(np.ascontiguousarray(torque, dtype=np.float64).reshape(nb * nb, 2))
    app = {
        "state": {tag: {"time": float(state["time"]), "iteration": [int(num), int(den)], "solution": solution}},
        "tasks": tasks if tasks is not None else new_tasks(),
        "time_series": {tag: series},
        "config": {"hydro": {tag: _ref_struct(hydro, tag)}, "model": cmodel, "mesh": _ref_struct(mesh, "mesh"),
                   "physics": _ref_struct(physics, "physics"), "control": _ref_struct(control or {}, "control")},
        "version": VERSION,
    }
    if ext:
        app[EXT] = ext
    return app


def unpack(app):
    """App tree -> dict(state, tasks, time_series, config, version), state as
    arrays in the boundary layout."""
    (tag, st), = app["state"].items()
    sol = st["solution"]
    keys = sorted(sol)
    nb = max(k[0] for k in keys) + 1
    if len(keys) != nb * nb or any(k != (b // nb, b % nb) for b, k in zip(range(nb * nb), keys)):
        raise ValueError("checkpoint solution is not a full square of blocks")
    first = sol[keys[0]]["conserved"]
    bs = int(first["dim"][0])
    data0 = np.asarray(first["data"], dtype=np.float64)
    nv = data0.shape[-1]
    u = np.empty((nb * nb, bs, bs, nv))
    ist = np.empty((nb * nb, 7, nv))
    oec = np.empty((nb * nb, 7, 4))
    for b, k in enumerate(keys):
        blk = sol[k]
        c = blk["conserved"]
        if int(c.get("v", 1)) != 1 or [int(d) for d in c["dim"]] != [bs, bs]:
            raise ValueError(f"block {k}: unexpected array header")
        u[b] = np.asarray(c["data"], dtype=np.float64).reshape(bs, bs, nv)
        ist[b] = _unitems(blk["integrated_source_terms"], nv)
        oec[b] = _unitems(blk["orbital_elements_change"], 4)
    num, den = st["iteration"]
    series = []
    for s in app.get("time_series", {}).get(tag, []):
        e = {"time": float(s["time"]), "ist": _unitems(s["integrated_source_terms"], nv),
             "oec": _unitems(s["orbital_elements_change"], 4)}
        if "torque" in s:
            e["torque"] = np.asarray(s["torque"], dtype=np.float64)
        series.append(e)
    ext = app.get(EXT) or {}
    config = app.get("config")
    model = (config or {}).get("model")
    if "model" in ext:
        model = ext["model"]
    if config is not None and "cooling_beta" in ext and "euler" in config.get("hydro", {}):
        config = {**config, "hydro": {"euler": {**config["hydro"]["euler"], "cooling_beta": ext["cooling_beta"]}}}
    torque = None
    if "torque" in ext:
        torque = np.asarray(ext["torque"], dtype=np.float64).reshape(nb * nb, 2)
    return {"state": {"time": float(st["time"]), "iteration": (int(num), int(den)), "conserved": u,
                      "ist": ist, "oec": oec, "torque": torque},
            "hydro": tag, "num_blocks": nb, "block_size": bs, "model": model,
            "tasks": app.get("tasks"), "time_series": series, "config": config,
            "version": app.get("version")}


def restore(solver, ckpt):
    """Put an unpacked checkpoint's state (unpack()["state"]) into a solver,
    with its torque totals when the checkpoint holds them."""
    st = ckpt["state"]
    solver.set_state(st["conserved"], time=st["time"], iteration=st["iteration"], ist=st["ist"], oec=st["oec"],
                     torque=st.get("torque"))


def write(path, state, mesh, hydro, physics, **kw):
    with open(path, "wb") as f:
        f.write(cbor.dumps(pack(state, mesh, hydro, physics, **kw)))


def read(path):
    with open(path, "rb") as f:
        return unpack(cbor.loads(f.read()))
