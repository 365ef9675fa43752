"""The run loop (src/main.rs:99-210) on the GPU path: time-series samples
(State::time_series_sample), device snapshots, checkpoint write and restart,
and the fallback stack with safe mode under a forced failure.  Bars:
bitwise where the computation is the same (restart, snapshot restore, a run
rewound to t = 0 in safe mode against a run in safe mode from the start)."""
import math

import numpy as np
import pytest

from cdc_amd import Snapshot, Solver, abi
from cdc_amd import checkpoint as ckpt
from cdc_amd.driver import Control, Runner, new_tasks
from tests import helpers as H

pytestmark = pytest.mark.gpu
ORBIT = 2.0 * math.pi


@pytest.fixture(scope="module", autouse=True)
def _device():
    lib = abi.load()
    assert lib.cdc_device_count() >= 1, "no HIP device: the HIP path must run on a GPU"


def states_equal(a, b):
    for k in ("conserved", "ist", "oec"):
        np.testing.assert_array_equal(a[k], b[k])
    assert a["time"] == b["time"] and a["iteration"] == b["iteration"]


def test_time_series_sample_sums_blocks():
    mesh, hydro, ph = H.small("euler", nb=3, bs=16, R=4.0)
    g = Solver(mesh, hydro, ph, H.initial(mesh, hydro, ph))
    g.advance(3)
    s, ts = g.get_state(), g.time_series_sample()
    assert ts["time"] == s["time"]
    np.testing.assert_allclose(ts["ist"], s["ist"].sum(axis=0), rtol=1e-13, atol=1e-13 * np.abs(s["ist"]).max())
    np.testing.assert_allclose(ts["oec"], s["oec"].sum(axis=0), rtol=1e-13, atol=1e-13 * np.abs(s["oec"]).max())
    np.testing.assert_allclose(ts["torque"], g.torque().sum(axis=0), rtol=1e-13)


def test_snapshot_restore_is_exact():
    mesh, hydro, ph = H.small("iso", nb=2, bs=16, R=4.0)
    g = Solver(mesh, hydro, ph, H.initial(mesh, hydro, ph))
    g.advance(2)
    snap = Snapshot(g)
    g.save(snap)
    before, tq = g.get_state(), g.torque()
    assert snap.time() == before["time"]
    g.advance(3)
    g.restore(snap)
    states_equal(g.get_state(), before)
    np.testing.assert_array_equal(g.torque(), tq)
    # and the restored solver continues exactly as the original would
    h = Solver(mesh, hydro, ph, H.initial(mesh, hydro, ph))
    h.advance(2)
    g.advance(4)
    h.advance(4)
    states_equal(g.get_state(), h.get_state())


def test_checkpoints_and_restart_continue_bitwise(tmp_path):
    mesh, hydro, ph = H.small("iso", nb=2, bs=16, R=4.0)
    u0 = H.initial(mesh, hydro, ph)
    ctl = Control(num_orbits=0.04, checkpoint_interval=0.02, time_series_interval=0.01, fold=2,
                  output_directory=str(tmp_path / "a"))
    logs = []
    full = Runner(Solver(mesh, hydro, ph, u0), ph, ctl, log=logs.append).run()
    # one time-series sample per side-effects call at most (each advance is ~0.025 orbit)
    assert len(full.checkpoints) >= 3 and len(full.time_series) >= 3
    # restart from the middle checkpoint and finish
    r = ckpt.read(full.checkpoints[1])
    g = Solver(mesh, hydro, ph, u0)
    ckpt.restore(g, r)
    tasks = new_tasks({k: v for k, v in r["tasks"].items()})
    series = list(r["time_series"])
    assert all("torque" in s for s in series)
    ctl2 = Control(**{**vars(ctl), "output_directory": str(tmp_path / "b")})
    rest = Runner(g, ph, ctl2, tasks=tasks, time_series=series, log=logs.append).run()
    a, b = full.solver.get_state(), rest.solver.get_state()
    np.testing.assert_array_equal(a["conserved"], b["conserved"])
    np.testing.assert_array_equal(a["ist"], b["ist"])
    assert a["time"] == b["time"] and a["iteration"] == b["iteration"]
    # the per-block torque totals carry on through the restart
    np.testing.assert_array_equal(full.solver.torque(), rest.solver.torque())
    # as in the reference, the restarted loop's first side effects may take a
    # second sample at the checkpoint's time when the task's next_time lags
    # behind (RecurringTask::advance adds the interval to next_time)
    assert sorted({s["time"] for s in rest.time_series}) == [s["time"] for s in full.time_series]
    assert rest.tasks["write_checkpoint"].count == full.tasks["write_checkpoint"].count
    # ... and the torque and accretion series (and the torque rates taken from
    # them) equal the uninterrupted run's
    from cdc_amd import timeseries as TS
    uniq = {}
    for s in rest.time_series:
        uniq.setdefault(s["time"], s)
    rs = TS.extract([uniq[t] for t in sorted(uniq)])
    fs = TS.extract(full.time_series)
    np.testing.assert_array_equal(rs["torque"], fs["torque"])
    np.testing.assert_array_equal(rs["ist"], fs["ist"])
    np.testing.assert_array_equal(TS.rates(rs)[2], TS.rates(fs)[2])


def test_fallback_rewinds_and_proceeds_in_safe_mode(tmp_path):
    """CFL 1.0 fails on the first advance (NegativeDensity); the loop rewinds
    to the oldest fallback state and continues with the safe physics, so the
    run equals one made with the safe physics from the start."""
    mesh, hydro, ph = H.small("iso", nb=2, bs=16, R=4.0, cfl=1.0, safe_cfl=0.2, safe_mode_duration=10.0)
    u0 = H.initial(mesh, hydro, ph)
    ctl = Control(num_orbits=0.01, checkpoint_interval=10.0, fold=2, fallback_stack_size=3,
                  output_directory=str(tmp_path))
    logs = []
    r = Runner(Solver(mesh, hydro, ph, u0), ph, ctl, log=logs.append).run()
    assert sum("rewind to orbit 0.00000" in m for m in logs) == 1, logs
    assert r.solver.time() >= 0.01 * ORBIT
    safe = ph.safe()
    ref = Solver(mesh, hydro, safe, u0)
    while ref.time() < 0.01 * ORBIT:
        ref.advance(2)
    states_equal(r.solver.get_state(), ref.get_state())


def test_failure_without_fallback_writes_fail_checkpoint(tmp_path):
    mesh, hydro, ph = H.small("iso", nb=2, bs=16, R=4.0, cfl=1.0)
    u0 = H.initial(mesh, hydro, ph)
    ctl = Control(num_orbits=0.01, checkpoint_interval=10.0, fold=2, output_directory=str(tmp_path))
    from cdc_amd import HydroError
    with pytest.raises(HydroError):
        Runner(Solver(mesh, hydro, ph, u0), ph, ctl, log=lambda m: None).run()
    r = ckpt.read(str(tmp_path / "chkpt.fail.cbor"))
    np.testing.assert_array_equal(r["state"]["conserved"], u0)
