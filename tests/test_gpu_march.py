"""The column-marching stage kernel's strip-edge arrays across everything
that changes a state outside its own stages: set_state, a snapshot restore,
stages of the tiled kernel, an ensemble step and the corrected advance of the
learned closure.  Each marching stage reads the strip-edge columns of its
input from an array that the stage producing that input wrote, and a state
no marching stage produced gets its array built first; a stale array would
put old values at every strip edge.  Bars as in test_gpu_parity."""
import numpy as np
import pytest

from cdc_amd import Ensemble, Snapshot, Solver, abi
from cdc_amd.closure import Closure, num_params
from tests import helpers as H
from tests.test_gpu_parity import check_states, pair

pytestmark = pytest.mark.gpu

STRIPS = (32, 64, 64)


@pytest.fixture(scope="module", autouse=True)
def _device():
    lib = abi.load()
    assert lib.cdc_device_count() >= 1, "no HIP device: the HIP path must run on a GPU"


def bumped(u, bs):
    """u with the density raised by 1% in the columns on both sides of every
    strip edge (a state the edge array of u does not describe)."""
    v = u.copy()
    for c in (0, 1, bs - 2, bs - 1):
        v[:, :, c, 0] *= 1.01
        v[:, :, c, 1:] *= 1.01
    return v


def test_edge_arrays_follow_the_committed_state():
    """The marching kernel against the oracle after set_state, after a
    snapshot restore and after switching to the tiled kernel and back."""
    mesh, hydro, ph = H.small("euler", nb=2, bs=64, R=4.0, rk_order="RK2")
    u0 = H.initial(mesh, hydro, ph)
    g, o = pair(mesh, hydro, ph, u0)
    g.set_kernel(*STRIPS)
    for _ in range(2):
        assert g.advance(1) == pytest.approx(o.advance(1), rel=1e-12)
    check_states(g, o)

    # set_state with strip-edge columns that differ from the last result
    so = o.get_state()
    u = bumped(so["conserved"], mesh.block_size)
    for s in (g, o):    # the torque totals start again from zero on both
        s.set_state(u, time=so["time"], iteration=so["iteration"], ist=so["ist"], oec=so["oec"])
    assert g.advance(1) == pytest.approx(o.advance(1), rel=1e-12)
    check_states(g, o)

    # a snapshot of this state, two steps on, then back to it
    snap = Snapshot(g)
    g.save(snap)
    g.advance(1)
    g.advance(1)
    g.restore(snap)
    snap.close()
    assert g.advance(1) == pytest.approx(o.advance(1), rel=1e-12)
    check_states(g, o)

    # tiled stages (which leave the edge arrays alone), then strips again
    g.set_kernel(16, 32, 256)
    assert g.advance(1) == pytest.approx(o.advance(1), rel=1e-12)
    g.set_kernel(*STRIPS)
    for _ in range(2):
        assert g.advance(1) == pytest.approx(o.advance(1), rel=1e-12)
    check_states(g, o)


ENSEMBLES = {
    # constant nu, constant nu, alpha viscosity
    "iso": [("iso", dict(rk_order="RK2")),
            ("iso", dict(rk_order="RK2", binary_eccentricity=0.3, binary_mass_ratio=0.5)),
            ("iso", dict(rk_order="RK2", viscosity={"alpha": 0.05}))],
    # T^4 cooling with alpha viscosity, beta cooling, T^4 cooling
    "euler": [("euler", dict(rk_order="RK3", viscosity={"alpha": 0.1})),
              ("euler_beta", dict(rk_order="RK3", binary_eccentricity=0.4)),
              ("euler", dict(rk_order="RK3", binary_mass_ratio=0.6))],
}


@pytest.mark.parametrize("which", sorted(ENSEMBLES))
def test_ensemble_of_marching_members(which):
    """Three 128^2 runs advanced as one ensemble with the strips, their
    cooling and viscosity laws mixed: each member bitwise equal to its solo
    run, the first against the oracle; one member's state is replaced
    between steps."""
    setups = [H.small(kind, nb=2, bs=64, R=4.0, **kw) for kind, kw in ENSEMBLES[which]]
    inits = [H.initial(m, h, p) for m, h, p in setups]
    batch = [Solver(m, h, p, u) for (m, h, p), u in zip(setups, inits)]
    solo = [Solver(m, h, p, u) for (m, h, p), u in zip(setups, inits)]
    for s in batch + solo:
        s.set_kernel(*STRIPS)
    o = H.oracle_for(*setups[0], inits[0])
    ens = Ensemble(batch)
    assert tuple(ens.get_tile()) == STRIPS[:2]
    for k in range(4):
        if k == 2:
            st = solo[1].get_state()
            u = bumped(st["conserved"], 64)
            for s in (batch[1], solo[1]):
                s.set_state(u, time=st["time"], iteration=st["iteration"], ist=st["ist"], oec=st["oec"],
                            torque=solo[1].torque())
        dt, codes, _ = ens.advance()
        assert np.all(codes == 0)
        for i, s in enumerate(solo):
            assert s.advance(1) == dt[i]
        o.advance(1)
    for a, b in zip(batch, solo):
        sa, sb = a.get_state(), b.get_state()
        np.testing.assert_array_equal(sa["conserved"], sb["conserved"])
        np.testing.assert_array_equal(sa["ist"], sb["ist"])
    check_states(batch[0], o)
    ens.close()


def test_corrected_advance_with_strips():
    """The learned closure's corrected advance on a 128^2 grid with the
    strips: an all-zero net is bitwise the plain marching advance (capture
    and replay); a nonzero net's corrected state matches the same corrected
    steps with the tiled kernel (the corrections change the strip-edge
    columns after every step) to 1e-3 of the correction applied."""
    mesh, hydro, ph = H.small("iso", nb=2, bs=64, R=4.0, rk_order="RK2")
    u0 = H.initial(mesh, hydro, ph)
    N = mesh.num_blocks * mesh.block_size
    net = Closure(N, mesh.domain_radius)
    net.set_params(np.zeros(num_params(), dtype=np.float32))
    a, b = Solver(mesh, hydro, ph, u0), Solver(mesh, hydro, ph, u0)
    for s in (a, b):
        s.set_kernel(*STRIPS)
    for _ in range(2):
        assert net.advance(a, fold=3) == b.advance(3)
    np.testing.assert_array_equal(a.get_state()["conserved"], b.get_state()["conserved"])
    net.close()

    rng = np.random.default_rng(4)
    net = Closure(N, mesh.domain_radius, limiter=0.1)
    net.set_params((rng.normal(size=num_params()) * 0.05).astype(np.float32))
    net.set_scales(output_scale=[1e-2] * 3)
    strips, tiles, plain = (Solver(mesh, hydro, ph, u0) for _ in range(3))
    strips.set_kernel(*STRIPS)
    tiles.set_kernel(16, 32, 256)
    plain.set_kernel(*STRIPS)
    for _ in range(2):
        dt = net.advance(strips, fold=3)
        assert net.advance(tiles, fold=3) == pytest.approx(dt, rel=1e-9)
        plain.advance_dt(dt, 3)
    us, ut, up = (s.get_state()["conserved"] for s in (strips, tiles, plain))
    moved = float(np.max(np.abs(us - up)))
    err = float(np.max(np.abs(us - ut)))
    assert moved > 0.0
    assert err <= 1e-3 * moved, (err, moved)
    net.close()


@pytest.mark.parametrize("rows", [32, 16, 8])
@pytest.mark.parametrize("kind,kw", [("iso", dict(rk_order="RK2")),
                                     ("euler", dict(rk_order="RK3", viscosity={"alpha": 0.1},
                                                    binary_eccentricity=0.5, binary_mass_ratio=0.5))],
                         ids=["iso", "euler"])
def test_wide_blocks_alternating_segments(kind, kw, rows):
    """Blocks of 128^2: two 64-column strips across each block and 4, 8 or 16
    row segments (strips of 32, 16 or 8 rows) that march down and up in
    turn, against the oracle after every call."""
    mesh, hydro, ph = H.small(kind, nb=2, bs=128, R=4.0, **kw)
    u0 = H.initial(mesh, hydro, ph)
    g, o = pair(mesh, hydro, ph, u0)
    g.set_kernel(rows, 64, 64)
    assert g.get_kernel() == (rows, 64, 64)
    for step in range(4):
        fold = 1 if step < 2 else 3
        assert g.advance(fold) == pytest.approx(o.advance(fold), rel=1e-12)
        check_states(g, o)


def test_strips_rejected_on_grids_beyond_32_bit_offsets():
    """The strips address their arrays through 32-bit buffer offsets: on a
    grid whose arrays hold 2^32 bytes or more (gamma-law gas, 11,648^2 =
    182 x 182 blocks of 64^2: 4.34 GB per array) set_kernel rejects them,
    the autotuner picks a tile shape, and a tiled stage runs."""
    from cdc_amd import CdcError, Euler, Mesh, workloads
    mesh = Mesh(182, 64, 8.0)
    hydro = Euler(gamma_law_index=5.0 / 3.0, cooling_coefficient=0.0)
    ph = workloads.base_physics(sink_rate=0.0)   # gas at rest under the binary, no sinks
    N = 182 * 64
    assert 4 * N * N * 8 >= 2 ** 32
    u0 = np.zeros((182 * 182, 64, 64, 4))
    u0[..., 0] = 1.0
    u0[..., 3] = 1.0 / (5.0 / 3.0 - 1.0)
    g = Solver(mesh, hydro, ph, u0)
    try:
        with pytest.raises(CdcError):
            g.set_kernel(32, 64, 64)
        g.set_kernel(16, 32, 256)
        assert g.get_kernel() == (16, 32, 256)
        g.autotune(reps=1)
        assert g.get_kernel()[1:] != (64, 64)
        assert g.advance(1) > 0.0
    finally:
        g.close()
