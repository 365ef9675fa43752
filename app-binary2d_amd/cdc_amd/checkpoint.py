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
The layout is restated from those conventions; no reference checkpoint file
exists to pin it, so it is covered by round-trip tests.  The configuration's
`model` is written as given (a reference model, or this build's seeded
cavity under the variant name "seeded_cavity").
"""
import dataclasses

import numpy as np

from . import cbor

ITEMS = ("sink1", "sink2", "grav1", "grav2", "buffer", "cooling", "fake_mass")
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


def pack(state, mesh, hydro, physics, model=None, control=None, tasks=None, time_series=None):
    """The App tree for a state dict (time, iteration, conserved [block][i][j][v],
    ist [block][7][nv], oec [block][7][4])."""
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
    return {
        "state": {tag: {"time": float(state["time"]), "iteration": [int(num), int(den)], "solution": solution}},
        "tasks": tasks if tasks is not None else new_tasks(),
        "time_series": {tag: series},
        "config": {"hydro": {tag: _plain(hydro)}, "model": model or {}, "mesh": _plain(mesh),
                   "physics": _plain(physics), "control": control or {}},
        "version": VERSION,
    }


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
    return {"state": {"time": float(st["time"]), "iteration": (int(num), int(den)), "conserved": u,
                      "ist": ist, "oec": oec},
            "hydro": tag, "num_blocks": nb, "block_size": bs,
            "tasks": app.get("tasks"), "time_series": series, "config": app.get("config"),
            "version": app.get("version")}


def write(path, state, mesh, hydro, physics, **kw):
    with open(path, "wb") as f:
        f.write(cbor.dumps(pack(state, mesh, hydro, physics, **kw)))


def read(path):
    with open(path, "rb") as f:
        return unpack(cbor.loads(f.read()))
