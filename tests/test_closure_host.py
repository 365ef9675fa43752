"""Host logic of the learned-closure study, with a stand-in solver (no GPU):
run_to ends a run exactly at the window's end, covering the remainder with
equal plain steps no longer than the CFL step, and reports a collapsing
time step; step_beside advances the fine run over the coarse step's dt in
sub-steps within the fine run's own CFL step."""
import math

import pytest

from cdc_amd.closure import Stalled, run_to, step_beside


class Stand:
    """time(), advance(fold) at a CFL step (fold 0: report it only),
    advance_dt(dt, fold) at a given step; the CFL step may shrink per call."""

    def __init__(self, cfl, shrink=1.0):
        self.t, self.cfl, self.shrink = 0.0, cfl, shrink
        self.plain = []

    def time(self):
        return self.t

    def advance(self, fold=1):
        dt = self.cfl
        if fold:
            self.t += fold * dt
            self.cfl *= self.shrink
        return dt

    def advance_dt(self, dt, fold=1):
        assert dt > 0.0
        self.plain.append((dt, fold))
        for _ in range(fold):
            self.t += dt


@pytest.mark.parametrize("cfl,fold,t_end", [(0.013, 10, 1.0), (0.013, 1, 1.0), (0.25, 10, 0.1),
                                            (1e-3, 10, 2 * math.pi)])
def test_run_to_lands_on_the_window_end(cfl, fold, t_end):
    s = Stand(cfl)
    n, _ = run_to(lambda v: fold * v.advance(fold), s, t_end, fold=fold)
    assert s.t == pytest.approx(t_end, rel=1e-13, abs=0.0)
    # the remainder: one call of equal plain steps, none longer than the CFL step
    assert len(s.plain) <= 1
    for dt, k in s.plain:
        assert dt <= cfl * (1 + 1e-15) and k <= fold
    # every regular call fits before t_end
    regular = n - len(s.plain)
    assert regular * fold * cfl < t_end + 1e-12


def test_run_to_reports_a_collapsing_step():
    s = Stand(0.01, shrink=0.5)
    with pytest.raises(Stalled):
        run_to(lambda v: v.advance(1), s, 10.0)


@pytest.mark.parametrize("dtc,dtf", [(0.04, 0.01), (0.04, 0.0099), (0.01, 0.02), (0.03, 0.0071)])
def test_step_beside_sub_steps_within_the_fine_cfl(dtc, dtf):
    c, f = Stand(dtc), Stand(dtf)
    dt = step_beside(c, f)
    assert dt == dtc
    (sub, m), = f.plain
    assert m == max(1, math.ceil(dtc / dtf))
    assert sub <= dtf * (1 + 1e-15)
    assert f.t == pytest.approx(c.t, rel=1e-14)


def test_sink_rates_match_the_oracle_sink_kernel():
    """closure._sink_rates (the per-cell sink rates of the sink-sampling
    diagnostic) against the oracle's Physics::sink_kernel of both bodies at
    the oracle's orbital state, at every cell centre of a small mesh, for a
    circular and an eccentric binary, at a time where the bodies sit off the
    grid's symmetry lines."""
    import numpy as np
    from cdc_amd import Mesh, workloads
    from cdc_amd.closure import _sink_rates
    from oracle import oracle as O
    mesh = Mesh(2, 32, 1.0)
    for e, q in ((0.0, 1.0), (0.5, 0.3)):
        ph = workloads.base_physics(binary_eccentricity=e, binary_mass_ratio=q, sink_radius=0.2)
        t = 0.37
        k, dx = _sink_rates(mesh, ph, t)
        pc = ph.to_c()
        st = O.orbital_state(pc, t)
        N = mesh.num_blocks * mesh.block_size
        assert dx == pytest.approx(2.0 / N, rel=1e-15)
        x = -1.0 + dx * (np.arange(N) + 0.5)
        want = np.array([[O.sink_kernel(pc, x[i] - st[0, 1], x[j] - st[0, 2]) +
                          O.sink_kernel(pc, x[i] - st[1, 1], x[j] - st[1, 2]) for j in range(N)] for i in range(N)])
        assert np.count_nonzero(want) > 8
        np.testing.assert_allclose(k, want, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("kw", [dict(rollout_rounds=1), dict(loop_rounds=1), dict(net_seeds=(6,), rollout_rounds=1)])
def test_closure_study_refuses_rounds_it_cannot_run(kw):
    """Loop and rollout rounds start from states of the evaluated run, so with
    training runs of their own they would train on it; extra net seeds
    retrain the one-step pairs only.  Both are refused before any run is
    built."""
    from cdc_amd import workloads
    from cdc_amd.closure import learned_vs_fine
    w = workloads.get("iso2048")
    train = () if "net_seeds" in kw else (w,)
    with pytest.raises(ValueError):
        learned_vs_fine(w, train_runs=train, **kw)


def test_variant_fields_defaults():
    """A learned-run variant of 3 fields corrects all components every step;
    4 and 5 fields set the components and the correction interval."""
    from cdc_amd.closure import variant_fields
    assert variant_fields((0.02, 0.25, 0.05)) == (0.02, 0.25, 0.05, 0, 1)
    assert variant_fields((0.02, 0.25, 0.05, 2)) == (0.02, 0.25, 0.05, 2, 1)
    assert variant_fields((0.02, 0.25, 0.05, 0, 2)) == (0.02, 0.25, 0.05, 0, 2)
    with pytest.raises(ValueError):
        variant_fields((0.02, 0.25))
