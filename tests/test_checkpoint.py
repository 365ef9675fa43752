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
    assert tree["config"]["hydro"] == {ckpt.hydro_tag(hydro): ckpt._plain(hydro)}


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
