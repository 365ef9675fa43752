"""The ensemble path (cdc_ensemble_*): one launch per RK stage over many
runs.  Every member must end bitwise where advancing it alone
(cdc_solver_advance(member, 1)) leaves it, with the graph replay and with
direct launches, under an active mask, and when a member fails."""
import numpy as np
import pytest

from cdc_amd import Ensemble, HydroError, Solver, abi
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    lib = abi.load()
    assert lib.cdc_device_count() >= 1, "no HIP device: the HIP path must run on a GPU"


def members_iso():
    out = []
    for i, (nb, bs, kw) in enumerate([
            (2, 32, dict(binary_eccentricity=0.3, binary_mass_ratio=0.7)),
            (4, 16, dict(viscosity={"alpha": 0.1})),
            (1, 64, dict(sink_model="torque_free", binary_mass_ratio=0.4)),
            (3, 16, dict(binary_eccentricity=0.6, viscosity={"constant_nu": 3e-3}))]):
        mesh, hydro, ph = H.small("iso", nb=nb, bs=bs, R=4.0, rk_order="RK2", **kw)
        out.append((mesh, hydro, ph, H.initial(mesh, hydro, ph, seed=11 + i)))
    return out


def members_euler():
    out = []
    for i, (kind, nb, bs, kw) in enumerate([
            ("euler", 2, 32, dict(viscosity={"alpha": 0.1})),
            ("euler_beta", 2, 16, dict(binary_eccentricity=0.4))]):
        mesh, hydro, ph = H.small(kind, nb=nb, bs=bs, R=4.0, rk_order="RK3", **kw)
        out.append((mesh, hydro, ph, H.initial(mesh, hydro, ph, seed=21 + i)))
    return out


def build(specs, graph=True):
    solo = [Solver(*s) for s in specs]
    batch = [Solver(*s) for s in specs]
    ens = Ensemble(batch, graph=graph)
    tile = ens.get_tile()
    for g in solo:
        g.set_tile(*tile)
    return solo, batch, ens


def same_state(a, b):
    sa, sb = a.get_state(), b.get_state()
    assert sa["time"] == sb["time"] and sa["iteration"] == sb["iteration"]
    np.testing.assert_array_equal(sa["conserved"], sb["conserved"])
    np.testing.assert_array_equal(sa["ist"], sb["ist"])
    np.testing.assert_array_equal(sa["oec"], sb["oec"])


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "direct"])
@pytest.mark.parametrize("which", ["iso", "euler"])
def test_members_match_solo_bitwise(which, graph):
    specs = members_iso() if which == "iso" else members_euler()
    solo, batch, ens = build(specs, graph)
    for step in range(4):
        dt, codes, errs = ens.advance()
        assert np.all(codes == 0) and all(e is None for e in errs)
        for i, g in enumerate(solo):
            assert g.advance(1) == dt[i], f"member {i} step {step}: dt differs"
            same_state(g, batch[i])
    n, z = ens.stats()
    assert n == 4 * len(specs)
    assert z == 4 * sum(m.mesh.total_zones() for m in batch)


def test_active_mask_and_solo_steps_between_batches():
    specs = members_iso()
    solo, batch, ens = build(specs)
    ens.advance()
    for g in solo:
        g.advance(1)
    active = np.array([1, 0, 1, 1], dtype=np.int32)
    before = batch[1].get_state()
    dt, codes, _ = ens.advance(active)
    assert np.isnan(dt[1]) and np.all(codes == 0)
    after = batch[1].get_state()
    np.testing.assert_array_equal(before["conserved"], after["conserved"])
    for i in (0, 2, 3):
        solo[i].advance(1)
    # a member advanced on its own between batch steps
    batch[1].advance(1)
    solo[1].advance(1)
    for _ in range(2):
        ens.advance()
        for g in solo:
            g.advance(1)
    for g, b in zip(solo, batch):
        same_state(g, b)


def test_failing_member_rolls_back_alone():
    specs = members_iso()
    solo, batch, ens = build(specs)
    ens.advance()
    for g in solo:
        g.advance(1)
    bad = batch[2].get_state()
    u = bad["conserved"].copy()
    u[0, 5, 7, 0] = -2e-9
    for g in (solo[2], batch[2]):
        g.set_state(u, time=bad["time"], iteration=bad["iteration"], ist=bad["ist"], oec=bad["oec"])
    dt, codes, errs = ens.advance()
    assert codes[2] == abi.CDC_ERR_NEGATIVE_DENSITY and isinstance(errs[2], HydroError)
    assert [codes[i] for i in (0, 1, 3)] == [0, 0, 0]
    with pytest.raises(HydroError) as es:
        solo[2].advance(1)
    e, s = errs[2], es.value
    assert (e.kind, e.value, e.position, e.time) == (s.kind, s.value, s.position, s.time)
    np.testing.assert_array_equal(batch[2].get_state()["conserved"], u)
    for i in (0, 1, 3):
        solo[i].advance(1)
    for g, b in zip(solo, batch):
        same_state(g, b)
    # the failed member is repaired and rejoins
    u[0, 5, 7, 0] = 2e-9
    for g in (solo[2], batch[2]):
        g.set_state(u, time=bad["time"], iteration=bad["iteration"], ist=bad["ist"], oec=bad["oec"])
    ens.advance()
    for g in solo:
        g.advance(1)
    for g, b in zip(solo, batch):
        same_state(g, b)


def test_ensemble_rejects_mixed_members():
    a = members_iso()[0]
    b = members_euler()[0]
    with pytest.raises(Exception):
        Ensemble([Solver(*a), Solver(*b)])


def test_shared_member_tile_is_kept():
    specs = members_euler()
    batch = [Solver(*s) for s in specs]
    for g in batch:
        g.set_tile(8, 8)
    ens = Ensemble(batch)
    assert ens.get_tile() == (8, 8)
    solo = [Solver(*s) for s in specs]
    for g in solo:
        g.set_tile(8, 8)
    dt, codes, _ = ens.advance()
    assert np.all(codes == 0)
    for i, g in enumerate(solo):
        assert g.advance(1) == dt[i]
        same_state(g, batch[i])
    # a shape that does not divide every block size falls back to the default
    other = [Solver(*s) for s in specs]
    other[0].set_tile(32, 16)
    other[1].set_tile(16, 16)
    assert Ensemble(other).get_tile() != (32, 16)


def test_ensemble64_workload_to_half_an_orbit():
    """The 64-run parameter study (workloads.ensemble_members: 24 x 256^2,
    24 x 512^2 and 16 x 1024^2 isothermal RK2 runs with drawn e, q, Mach and
    nu, blocks of 64^2) advanced as one ensemble until every run has reached
    half an orbit.  The first run of each resolution matches the oracle step
    by step for its first 6 steps (check_states: state within 1e-12, dt,
    items, orbital-element changes, torques); every run then ends bitwise
    where advancing it alone leaves it, and closes its mass budget: the
    change of its total mass is the sum of its sink, buffer and fake-mass
    items to 1e-9 of the total mass.  (Not to rounding: each block evaluates
    its own faces with its own coordinates, src/scheme.rs:257-268, and with
    an asymmetric binary the isothermal sound speed -phi/M^2 differs between
    the periodic edge faces at x = -R and x = +R, so the edge fluxes do not
    cancel; measured 4e-11 of the total mass after 1390 steps.)"""
    import math
    from cdc_amd import workloads
    from tests.test_gpu_parity import check_states
    members = workloads.ensemble_members()
    assert len(members) == 64
    inits = [workloads.initial_state(m) for m in members]
    batch = [Solver(m["mesh"], m["hydro"], m["physics"], u) for m, u in zip(members, inits)]
    ens = Ensemble(batch)
    tile = ens.get_tile()
    first = {}
    for i, m in enumerate(members):
        first.setdefault(m["mesh"].num_blocks, i)
    assert sorted(first) == [4, 8, 16]
    oracles = {i: H.oracle_for(members[i]["mesh"], members[i]["hydro"], members[i]["physics"], inits[i])
               for i in first.values()}
    target = 0.5 * 2.0 * math.pi
    steps = np.zeros(len(batch), dtype=np.int64)
    k = 0
    while True:
        active = np.array([g.time() < target for g in batch], dtype=np.int32)
        if not active.any():
            break
        dt, codes, errs = ens.advance(active)
        assert np.all(codes == 0), [e for e in errs if e is not None][:1]
        steps += active
        if k < 6:
            for i, o in oracles.items():
                assert dt[i] == pytest.approx(o.advance(1), rel=1e-12)
                check_states(batch[i], o)
        k += 1
    for o in oracles.values():
        o.close()
    print(f"ensemble64: {k} ensemble steps, {int(steps.sum())} member steps")
    for i, (m, u0, g) in enumerate(zip(members, inits, batch)):
        solo = Solver(m["mesh"], m["hydro"], m["physics"], u0)
        solo.set_tile(*tile)
        for _ in range(int(steps[i])):
            solo.advance(1)
        same_state(solo, g)
        np.testing.assert_array_equal(solo.torque(), g.torque())
        solo.close()
        s = g.get_state()
        dA = m["mesh"].cell_spacing() ** 2
        dm = (s["conserved"][..., 0].sum() - u0[..., 0].sum()) * dA
        src = s["ist"][:, :, 0].sum()
        assert abs(dm - src) <= 1e-9 * u0[..., 0].sum() * dA, (i, dm, src)
