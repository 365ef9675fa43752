"""CBOR codec and the App checkpoint layout (no GPU): round trips of every
CBOR item kind, decoding of the narrower float widths a serde writer may
choose, the vectorized conserved-block path against the general decoder, and
a full checkpoint tree round trip."""
import math
import struct

import numpy as np
import pytest

from cdc_amd import Euler, Isothermal, Mesh, cbor, workloads
from cdc_amd import checkpoint as ckpt


def test_cbor_items_round_trip():
    x = {"i": [0, 23, 24, 255, 256, 65535, 65536, 2 ** 32, 2 ** 40, -1, -24, -25, -2 ** 40],
         "f": [0.0, -0.0, 1.5, math.pi, 1e300, -1e-300, math.inf],
         "s": ["", "a", "x" * 300], "b": [b"", b"\x00\xff" * 200], "z": [None, True, False],
         (3, 4): {"nested": [[1, 2], {"k": 2.0}]}}
    y = cbor.loads(cbor.dumps(x))
    assert y["i"] == x["i"] and y["s"] == x["s"] and y["b"] == x["b"] and y["z"] == x["z"]
    assert [struct.pack(">d", v) for v in y["f"]] == [struct.pack(">d", v) for v in x["f"]]
    assert y[(3, 4)] == {"nested": [[1, 2], {"k": 2.0}]}


def test_cbor_known_encodings():
    # RFC 8949 appendix A examples
    assert cbor.dumps(0) == bytes.fromhex("00")
    assert cbor.dumps(1000) == bytes.fromhex("1903e8")
    assert cbor.dumps(-1000) == bytes.fromhex("3903e7")
    assert cbor.dumps(1.1) == bytes.fromhex("fb3ff199999999999a")
    assert cbor.dumps("IETF") == bytes.fromhex("6449455446")
    assert cbor.dumps([1, [2, 3]]) == bytes.fromhex("8201820203")
    assert cbor.dumps({"a": 1}) == bytes.fromhex("a1616101")
    assert cbor.loads(bytes.fromhex("f93c00")) == 1.0          # half
    assert cbor.loads(bytes.fromhex("f97bff")) == 65504.0
    assert cbor.loads(bytes.fromhex("fa47c35000")) == 100000.0  # single
    assert cbor.loads(bytes.fromhex("c11a514b67b0")) == 1363896240   # tag ignored


def test_rows_fast_path_equals_general_decoder():
    a = np.random.default_rng(0).normal(size=(1000, 4))
    b = cbor.dumps(cbor.Rows(a))
    assert b == cbor.dumps(a.tolist())
    np.testing.assert_array_equal(cbor.loads(b), a)
    # a block written with narrower floats (as a serde writer may) takes the general path
    mixed = cbor.dumps([[1.0, 2.0]])[:-9] + bytes.fromhex("f94000")      # last item as half 2.0
    mixed = bytes([0x81, 0x82]) + bytes.fromhex("fb3ff0000000000000") + bytes.fromhex("f94000")
    assert cbor.loads(mixed) == [[1.0, 2.0]]


@pytest.mark.parametrize("hydro", [Isothermal(mach_number=10.0), Euler(gamma_law_index=5 / 3, cooling_coefficient=0.0)])
def test_checkpoint_tree_round_trip(tmp_path, hydro):
    mesh = Mesh(3, 8, 4.0)
    nv = hydro.nv
    rng = np.random.default_rng(1)
    st = {"time": 1.25, "iteration": (17, 3), "conserved": rng.normal(size=(9, 8, 8, nv)),
          "ist": rng.normal(size=(9, 7, nv)), "oec": rng.normal(size=(9, 7, 4))}
    series = [{"time": 0.5, "ist": rng.normal(size=(7, nv)), "oec": rng.normal(size=(7, 4))}]
    tasks = ckpt.new_tasks()
    tasks["write_checkpoint"] = {"count": 3, "next_time": 6.28}
    p = tmp_path / "chkpt.0003.cbor"
    ckpt.write(str(p), st, mesh, hydro, workloads.base_physics(), model={"seeded_cavity": {"seed": 1}},
               control={"num_orbits": 2.0}, tasks=tasks, time_series=series)
    r = ckpt.read(str(p))
    for k in ("conserved", "ist", "oec"):
        np.testing.assert_array_equal(r["state"][k], st[k])
    assert r["state"]["time"] == 1.25 and r["state"]["iteration"] == (17, 3)
    assert r["num_blocks"] == 3 and r["block_size"] == 8 and r["hydro"] == ckpt.hydro_tag(hydro)
    assert r["tasks"]["write_checkpoint"] == {"count": 3, "next_time": 6.28}
    np.testing.assert_array_equal(r["time_series"][0]["ist"], series[0]["ist"])
    # the tree follows serde's layout: externally tagged state, tuple block keys,
    # ndarray's {v, dim, data}
    tree = cbor.loads(p.read_bytes())
    blk = tree["state"][ckpt.hydro_tag(hydro)]["solution"][(1, 2)]
    assert blk["conserved"]["v"] == 1 and blk["conserved"]["dim"] == [8, 8]
    np.testing.assert_array_equal(blk["conserved"]["data"], st["conserved"][5].reshape(64, nv))
    assert list(blk["integrated_source_terms"]) == list(ckpt.ITEMS)
    assert r["model"] == {"seeded_cavity": {"seed": 1}}


# The reference's serde field names, restated here from the reference's
# structs (not taken from the module under test): every one of these structs
# except Control rejects unknown fields.
REF_MESH = {"num_blocks", "block_size", "domain_radius"}                         # src/mesh.rs:19-37
REF_ISO = {"mach_number", "density_floor"}                                       # src/physics.rs:255-265
REF_EULER = {"gamma_law_index", "cooling_coefficient", "pressure_floor",         # src/physics.rs:270-303
             "density_floor", "density_index", "cooling_slow_start"}
REF_PHYSICS = {"buffer_rate", "buffer_scale", "binary_eccentricity",             # src/physics.rs:121-217
               "binary_mass_ratio", "lambda_multiplier", "viscosity", "sink_model", "cfl", "plm",
               "plm_density_trigger", "fake_mass_threshold", "fake_mass_rate", "one_body", "mach_ceiling",
               "rk_order", "sink_radius", "sink_rate", "softening_length", "dt_min", "safe_mode_duration",
               "safe_cfl", "safe_plm", "safe_mach_ceiling", "safe_rk_order", "lambda", "nu", "alpha"}
REF_PHYSICS_REQUIRED = {"buffer_rate", "buffer_scale", "binary_eccentricity", "binary_mass_ratio", "cfl", "plm",
                        "rk_order", "sink_radius", "sink_rate"}
REF_CONTROL_REQUIRED = {"num_orbits", "checkpoint_interval", "fold", "output_directory"}   # src/app.rs:94-131
REF_MODELS = {"infinite_disk": {"mach_number", "surface_density"},                # src/app.rs:41-48,
              "infinite_alpha_disk": {"mach_number", "surface_density"},          # src/disks.rs:11-93
              "finite_disk": {"disk_mass", "disk_width", "disk_radius", "softening_length", "mach_number"},
              "residual_test": set()}


def _check_reference_config(cfg, hydro_tag):
    assert set(cfg) == {"hydro", "model", "mesh", "physics", "control"}            # src/app.rs:169-177
    assert set(cfg["hydro"]) == {hydro_tag}
    h = cfg["hydro"][hydro_tag]
    assert set(h) <= (REF_ISO if hydro_tag == "isothermal" else REF_EULER)
    assert {"gamma_law_index", "cooling_coefficient"} <= set(h) if hydro_tag == "euler" else "mach_number" in h
    assert set(cfg["mesh"]) == REF_MESH
    assert set(cfg["physics"]) <= REF_PHYSICS and REF_PHYSICS_REQUIRED <= set(cfg["physics"])
    (vk, vv), = cfg["physics"]["viscosity"].items()
    assert vk in ("constant_nu", "alpha") and isinstance(vv, float)
    assert cfg["physics"]["sink_model"] in ("impartial", "torque_free")
    assert cfg["physics"]["rk_order"] in ("RK1", "RK2", "RK3")
    assert REF_CONTROL_REQUIRED <= set(cfg["control"])
    (mk, mv), = cfg["model"].items()
    assert mk in REF_MODELS and set(mv) <= REF_MODELS[mk]
    # exactly one Mach number between hydro and model (InfiniteDiskModel::validate)
    if mk.startswith("infinite"):
        assert (mv.get("mach_number") is not None) != (h.get("mach_number") is not None)


@pytest.mark.parametrize("name", ["iso256", "euler2048"])
def test_checkpoint_config_has_only_reference_fields(tmp_path, name):
    """The written config is readable by the reference's deny_unknown_fields
    structs: every key is one of the reference's serde fields (so `lambda`,
    not `lambda_`, and no `cooling_beta` in euler), `model` is an AnyModel
    variant, and the extensions (beta cooling, the seeded cavity model,
    per-block torques) sit under the App-level extension key and come back."""
    w = workloads.get(name)
    mesh, hydro, ph = Mesh(2, 8, 8.0), w["hydro"], w["physics"]
    ph.lambda_ = 0.25
    nv = hydro.nv
    rng = np.random.default_rng(3)
    st = {"time": 0.5, "iteration": (3, 1), "conserved": rng.normal(size=(4, 8, 8, nv)),
          "ist": rng.normal(size=(4, 7, nv)), "oec": rng.normal(size=(4, 7, 4))}
    torque = rng.normal(size=(4, 2))
    model = {"seeded_cavity": {"seed": 3, "cavity_radius": 2.5}}
    p = tmp_path / "chkpt.0000.cbor"
    ckpt.write(str(p), st, mesh, hydro, ph, model=model, control={"num_orbits": 2.0, "fold": 10},
               torque=torque)
    tree = cbor.loads(p.read_bytes())
    tag = ckpt.hydro_tag(hydro)
    _check_reference_config(tree["config"], tag)
    assert tree["config"]["physics"]["lambda"] == 0.25
    assert set(tree) == {"state", "tasks", "time_series", "config", "version", "cdc_amd"}
    r = ckpt.read(str(p))
    assert r["model"] == model
    np.testing.assert_array_equal(r["state"]["torque"], torque)
    # the config reads back into this package's types with the extensions folded in
    from cdc_amd.config import Configuration
    cfg = Configuration.from_dict({**r["config"], "model": r["model"]})
    assert cfg.physics.lambda_ == 0.25
    assert getattr(cfg.hydro, "cooling_beta", None) == getattr(hydro, "cooling_beta", None)


def test_checkpoint_keeps_a_reference_model():
    mesh, hydro = Mesh(1, 4, 4.0), Isothermal(mach_number=10.0)
    st = {"time": 0.0, "iteration": (0, 1), "conserved": np.ones((1, 4, 4, 3)), "ist": np.zeros((1, 7, 3)),
          "oec": np.zeros((1, 7, 4))}
    model = {"finite_disk": {"disk_mass": 0.1, "disk_width": 1.5, "disk_radius": 3.0, "softening_length": 0.05,
                             "mach_number": None}}
    tree = ckpt.pack(st, mesh, hydro, workloads.base_physics(), model=model)
    _check_reference_config(tree["config"], "isothermal")
    assert tree["config"]["model"] == model and "cdc_amd" not in tree


def test_time_series_rates_and_checkpoint_round_trip(tmp_path):
    """Accretion and torque rates from samples with linear ramps are exact,
    and the samples (with the torque extension) survive a checkpoint."""
    from cdc_amd import timeseries as ts
    times = np.array([0.0, 0.5, 1.25, 2.0])
    samples = []
    for t in times:
        ist = np.zeros((7, 3))
        ist[0, 0] = -0.25 * t          # sink 1 accretes at 0.25
        ist[1, 0] = -0.75 * t          # sink 2 at 0.75
        samples.append({"time": t, "ist": ist, "oec": np.zeros((7, 4)), "torque": np.array([2.0 * t, -t])})
    s = ts.extract(samples)
    tm, mdot, tau = ts.rates(s)
    np.testing.assert_allclose(tm, [0.25, 0.875, 1.625])
    np.testing.assert_allclose(mdot, [[0.25, 0.75]] * 3, rtol=1e-15)
    np.testing.assert_allclose(tau, [[2.0, -1.0]] * 3, rtol=1e-15)
    np.testing.assert_allclose(ts.accreted_mass(s)[-1], [0.5, 1.5])
    mesh = Mesh(1, 8, 4.0)
    st = {"time": 2.0, "iteration": (4, 1), "conserved": np.ones((1, 8, 8, 3)), "ist": np.zeros((1, 7, 3)),
          "oec": np.zeros((1, 7, 4))}
    p = tmp_path / "chkpt.0001.cbor"
    ckpt.write(str(p), st, mesh, Isothermal(mach_number=10.0), workloads.base_physics(), time_series=samples)
    back = ts.from_checkpoint(str(p))
    np.testing.assert_array_equal(back["torque"], s["torque"])
    np.testing.assert_array_equal(back["ist"], s["ist"])
    ts.main([str(p), "--output", str(tmp_path / "series.npz")])
    z = np.load(tmp_path / "series.npz")
    np.testing.assert_allclose(z["mdot"], mdot)
