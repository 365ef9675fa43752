"""Learned flux closure for coarse runs (cdc_closure_*), and the pipeline
that trains it on a down-sampled fine run and compares coarse runs with and
without it against the fine run.  No counterpart in the reference."""
import ctypes
import math
import time
from ctypes import byref, c_double, c_int32, c_void_p

import numpy as np

from . import abi
from .solver import CdcError, HydroError, _check, _ptr

NUM_PARAMS = None


def num_params():
    global NUM_PARAMS
    if NUM_PARAMS is None:
        NUM_PARAMS = int(abi.load().cdc_closure_num_params())
    return NUM_PARAMS


class Closure:
    """Net, Adam state and training samples on the GPU for one coarse grid."""

    def __init__(self, grid, domain_radius, max_samples=0, seed=1, feature_scale=(0.25, 0.5, 0.5),
                 output_scale=(1.0, 1.0, 1.0), loss_weight=(1.0, 1.0, 1.0), init_output_gain=1.0, limiter=0.0,
                 mask_density=0.0):
        self.lib = abi.load()
        c = abi.CClosureConfig()
        c.grid, c.max_samples, c.seed, c.domain_radius = int(grid), int(max_samples), int(seed), float(domain_radius)
        c.feature_scale[:] = list(feature_scale)
        c.output_scale[:] = list(output_scale)
        c.loss_weight[:] = list(loss_weight)
        c.init_output_gain = float(init_output_gain)
        c.limiter = float(limiter)
        c.mask_density = float(mask_density)
        h = c_void_p()
        _check(self.lib, self.lib.cdc_closure_create(byref(c), byref(h)))
        self.handle = h
        self.grid = int(grid)
        self.capacity = int(max_samples)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.cdc_closure_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def config(self):
        c = abi.CClosureConfig()
        _check(self.lib, self.lib.cdc_closure_get_config(self.handle, byref(c)))
        return c

    def params(self):
        out = np.empty(num_params(), dtype=np.float32)
        _check(self.lib, self.lib.cdc_closure_get_params(self.handle, _ptr(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32).reshape(num_params())
        _check(self.lib, self.lib.cdc_closure_set_params(self.handle, _ptr(p)))

    def set_scales(self, output_scale=None, loss_weight=None):
        o = None if output_scale is None else np.ascontiguousarray(output_scale, dtype=np.float64)
        w = None if loss_weight is None else np.ascontiguousarray(loss_weight, dtype=np.float64)
        _check(self.lib, self.lib.cdc_closure_set_scales(self.handle, None if o is None else _ptr(o),
                                                         None if w is None else _ptr(w)))

    def set_interval(self, k):
        """Correct every k-th step of an advance: the net on a group's
        starting state, the correction over the group's k steps."""
        _check(self.lib, self.lib.cdc_closure_set_interval(self.handle, int(k)))

    def features(self, solver):
        out = np.empty((self.grid * self.grid, 4), dtype=np.float32)
        _check(self.lib, self.lib.cdc_closure_features(self.handle, solver.handle, _ptr(out)))
        return out.reshape(self.grid, self.grid, 4)

    def forward(self, solver):
        out = np.empty((self.grid * self.grid, 6), dtype=np.float32)
        _check(self.lib, self.lib.cdc_closure_forward(self.handle, solver.handle, _ptr(out)))
        return out.reshape(self.grid, self.grid, 6)

    def add_sample(self, base, plain, truth, dt):
        _check(self.lib, self.lib.cdc_closure_add_sample(self.handle, base.handle, plain.handle, truth.handle,
                                                         float(dt)))

    def samples(self):
        n = c_int32()
        ms = np.zeros(3)
        _check(self.lib, self.lib.cdc_closure_num_samples(self.handle, byref(n), _ptr(ms)))
        return n.value, ms

    def clear_samples(self):
        _check(self.lib, self.lib.cdc_closure_clear_samples(self.handle))

    def sample_grad(self, k):
        g = np.empty(num_params(), dtype=np.float32)
        l = c_double()
        _check(self.lib, self.lib.cdc_closure_sample_grad(self.handle, int(k), byref(l), _ptr(g)))
        return l.value, g

    def train(self, steps, lr=1e-3, seed=0):
        out = np.empty(max(1, steps), dtype=np.float32)
        _check(self.lib, self.lib.cdc_closure_train(self.handle, int(steps), float(lr), int(seed), _ptr(out)))
        return out[:steps]

    def advance(self, solver, fold=1):
        """`fold` corrected coarse steps at one dt (one replayed graph, one
        host synchronisation); returns dt, raises HydroError."""
        dt = c_double()
        err = abi.CError()
        rc = self.lib.cdc_closure_advance_fold(self.handle, solver.handle, int(fold), byref(dt), byref(err))
        if rc in (abi.CDC_ERR_NEGATIVE_DENSITY, abi.CDC_ERR_NEGATIVE_PRESSURE, abi.CDC_ERR_UNBOUND_ORBIT):
            raise HydroError(err)
        _check(self.lib, rc)
        return dt.value


ORBIT = 2.0 * math.pi


def accreted(solver):
    """Mass taken by both sinks so far: minus the sum over blocks of the
    sink items' density component of integrated_source_terms."""
    ist = solver.get_state()["ist"]
    return -float(ist[:, 0, 0].sum() + ist[:, 1, 0].sum())


class Stalled(RuntimeError):
    """The time step collapsed (a run that would never reach its end)."""


def run_to(step, solver, t_end, log=None, every=2000, min_dt_ratio=1e-2, fold=1):
    """Advance with `step(solver)` (which returns the time advanced, at most
    `fold` CFL steps) until the solver's time is exactly t_end: once the next
    call could pass t_end, the rest is covered by equal plain steps no longer
    than the CFL step (advance_dt).  Raises Stalled when a step advances less
    than min_dt_ratio of the first one.  Returns (calls, wall seconds)."""
    n = 0
    t0 = time.perf_counter()
    first = None
    while True:
        t = solver.time()
        cfl = solver.advance(0)        # the CFL step of the committed state; nothing advanced
        if t + fold * cfl >= t_end:
            break
        d = step(solver)
        first = d if first is None else first
        if d < min_dt_ratio * first:
            raise Stalled(f"time step collapsed to {d:.3e} (first {first:.3e}) at t = {t / ORBIT:.4f} orbits")
        n += 1
        if log and n % every == 0:
            log(f"  t = {t / ORBIT:.3f} orbits after {n} steps")
    rem = t_end - solver.time()
    if rem > 0.0:
        k = max(1, math.ceil(rem / cfl))
        solver.advance_dt(rem / k, k)
        n += 1
    return n, time.perf_counter() - t0


def heldout_one_step(fine, net, coarse_mesh, hydro, physics, cu0, pairs=16, gap_steps=40, burn_in=0):
    """The trained closure against the plain coarse step on states it was not
    trained on: from `pairs` down-sampled states of the fine run, gap_steps
    fine steps apart, one plain and one corrected coarse step (both start from
    the same state, so both take the same CFL dt), each compared with the fine
    run advanced over that dt and down-sampled (a state from which either
    coarse step fails is skipped and counted).  Returns the L1 errors per
    conserved field (Sigma, Sigma vx, Sigma vy) summed over the pairs, plain
    and learned, and their ratio (below 1: the closure gains).  With
    burn_in > 0 both steps start from a coarse run stepped burn_in times
    beside the fine run, and the errors are those of the increments (the
    step's change against the down-sampled fine run's change)."""
    from .solver import Solver
    plain, corr, truth = (Solver(coarse_mesh, hydro, physics, cu0) for _ in range(3))
    e_plain, e_learned = np.zeros(3), np.zeros(3)
    skipped = 0
    try:
        for _ in range(pairs):
            fine.advance(gap_steps)
            plain.resample_from(fine)
            for _ in range(burn_in):
                step_beside(plain, fine)
            truth.resample_from(fine)
            t0 = truth.get_state()["conserved"]
            ub = plain.get_state()["conserved"]
            corr.resample_from(plain)
            try:
                dt = plain.advance(1)
                dtl = net.advance(corr)
            except HydroError:
                skipped += 1       # a coarse step failed from this state: no comparison
                continue
            if abs(dtl - dt) > 1e-12 * dt:
                raise RuntimeError(f"corrected step took dt {dtl!r}, plain step {dt!r}")
            m = max(1, math.ceil(dt / fine.advance(0)))
            fine.advance_dt(dt / m, m)
            truth.resample_from(fine)
            # the state the fine run's increment carries the base to
            ut = ub + (truth.get_state()["conserved"] - t0)
            up = plain.get_state()["conserved"]
            ul = corr.get_state()["conserved"]
            for v in range(3):
                e_plain[v] += np.abs(up[..., v] - ut[..., v]).sum()
                e_learned[v] += np.abs(ul[..., v] - ut[..., v]).sum()
    finally:
        for s in (plain, corr, truth):
            s.close()
    return {"pairs": pairs - skipped, "skipped": skipped, "gap_fine_steps": gap_steps, "burn_in": burn_in,
            "l1_plain": e_plain.tolist(),
            "l1_learned": e_learned.tolist(), "ratio": (e_learned / e_plain).tolist()}


def step_beside(coarse, fine):
    """One plain CFL step of the coarse solver, and the fine solver advanced
    over the same time in equal sub-steps no longer than its own CFL step.
    Returns the coarse step's dt."""
    dt = coarse.advance(1)
    m = max(1, math.ceil(dt / fine.advance(0)))
    fine.advance_dt(dt / m, m)
    return dt


def loop_pairs(net, fl, snap, coarse_mesh, hydro, physics, cu0, steps=200, every=10):
    """Model-in-the-loop training pairs: the fine solver fl restarts from the
    device snapshot snap, the coarse state starts as its down-sampling and is
    stepped `steps` times with the closure (net.advance) while fl advances
    over the same time steps (each in equal sub-steps no longer than its CFL
    step); every `every` steps the pair is (the corrected run's state, its
    plain coarse step, fl down-sampled at the step's end).  The targets thus
    hold the drift of the corrected run from the fine trajectory as well as
    the one-step error.  Returns (pairs added, 1 if the corrected run failed
    before the end else 0)."""
    from .solver import Solver
    c, base, plain, truth = (Solver(coarse_mesh, hydro, physics, cu0) for _ in range(4))
    added, failed = 0, 0
    try:
        fl.restore(snap)
        c.resample_from(fl)
        for k in range(steps):
            take = k % every == 0 and net.samples()[0] < net.capacity
            if take:
                base.resample_from(c)
                plain.resample_from(c)
                dt = plain.advance(1)
            try:
                dtc = net.advance(c)
            except HydroError:
                failed = 1
                break
            if take and abs(dtc - dt) > 1e-12 * dt:
                raise RuntimeError(f"corrected step took dt {dtc!r}, plain step {dt!r}")
            m = max(1, math.ceil(dtc / fl.advance(0)))
            fl.advance_dt(dtc / m, m)
            if take:
                truth.resample_from(fl)
                net.add_sample(base, plain, truth, dtc)
                added += 1
    finally:
        for s in (c, base, plain, truth):
            s.close()
    return added, failed


def _global(u, nb, bs):
    """[nb*nb][bs][bs][nv] block layout -> [N][N][nv] (i-major)."""
    nv = u.shape[-1]
    return u.reshape(nb, nb, bs, bs, nv).transpose(0, 2, 1, 3, 4).reshape(nb * bs, nb * bs, nv)


def _blocks(u, nb, bs):
    nv = u.shape[-1]
    return u.reshape(nb, bs, nb, bs, nv).transpose(0, 2, 1, 3, 4).reshape(nb * nb, bs, bs, nv)


def flux_divergence(Y, scale, dx):
    """The correction rate r_v = -s_v ((Y_v - Y_v[i-1]) / dx + (Y_3+v - Y_3+v[j-1]) / dx)
    of the net output Y ([N][N][6], i-major), v = 0..2: what the training
    loss compares with its target (no limiter, no mask)."""
    r = np.empty(Y.shape[:2] + (3,))
    for v in range(3):
        r[..., v] = -scale[v] * ((Y[..., v] - np.roll(Y[..., v], 1, axis=0)) / dx +
                                 (Y[..., 3 + v] - np.roll(Y[..., 3 + v], 1, axis=1)) / dx)
    return r


def _box(a, w):
    """Periodic w x w box mean of a [N][N][...] field (w = 1: unchanged)."""
    if w <= 1:
        return a
    out = np.zeros_like(a)
    for di in range(w):
        for dj in range(w):
            out += np.roll(np.roll(a, di - w // 2, axis=0), dj - w // 2, axis=1)
    return out / (w * w)


def rollout_pairs(net, fl, snap, coarse_mesh, hydro, physics, cu0, scale, steps=200, every=20, min_fit=10,
                  gain=1.0, smooth=1):
    """Training pairs from a rollout (a posteriori) objective.  The fine
    solver fl restarts from the device snapshot snap; the coarse state starts
    as its down-sampling and takes `steps` corrected steps while fl advances
    over the same times; e_j is the corrected run's error against fl
    down-sampled after step j.  The rollout loss sum_j |e_j|^2 / 2 has the
    gradient dt sum_{j>m} e_j with respect to the correction rate of step m
    when the plain step's Jacobian is taken as the identity; the target of
    the pair at step m (every `every` steps, with at least min_fit steps
    after it) is that rate minus b_m, the least-squares growth rate of the
    error over the steps after m (e_j - e_m against t_j - t_m), so that a
    correction that leaves a steady error growth behind is moved by it.
    `gain` scales b (a damped step) and `smooth` box-filters it over smooth x
    smooth cells (only the error growth at that scale is targeted).
    Returns (pairs added, mean relative L1 error of Sigma over the rollout,
    1 if the corrected run failed before the end else 0)."""
    from .solver import Solver
    nb, bs = coarse_mesh.num_blocks, coarse_mesh.block_size
    N = nb * bs
    dx = 2.0 * coarse_mesh.domain_radius / N
    c, base, plain, truth = (Solver(coarse_mesh, hydro, physics, cu0) for _ in range(4))
    open_ = []          # per pending pair: [state dict, dt, t_m, e_m, S1, S2]
    added, failed, l1 = 0, 0, []
    try:
        fl.restore(snap)
        c.resample_from(fl)
        e_prev = np.zeros((N, N, 3))
        for k in range(steps):
            if k % every == 0 and steps - k >= min_fit and net.samples()[0] + len(open_) < net.capacity:
                open_.append([c.get_state(), None, c.time(), e_prev.copy(), np.zeros((N, N, 3)), 0.0])
            try:
                dtc = net.advance(c)
            except HydroError:
                failed = 1
                break
            for o in open_:
                if o[1] is None:
                    o[1] = dtc
            m = max(1, math.ceil(dtc / fl.advance(0)))
            fl.advance_dt(dtc / m, m)
            truth.resample_from(fl)
            ut = _global(truth.get_state()["conserved"], nb, bs)
            uc = _global(c.get_state()["conserved"], nb, bs)
            e = uc[..., :3] - ut[..., :3]
            l1.append(float(np.abs(e[..., 0]).sum() / np.abs(ut[..., 0]).sum()))
            tj = c.time()
            for o in open_:
                d = tj - o[2]
                o[4] += d * (e - o[3])
                o[5] += d * d
            e_prev = e
        for st, dt, _, _, S1, S2 in open_:
            if dt is None or S2 <= 0.0:
                continue
            b = gain * _box(S1 / S2, smooth)
            base.set_state(st["conserved"], time=st["time"], iteration=st["iteration"], ist=st["ist"],
                           oec=st["oec"])
            plain.set_state(st["conserved"], time=st["time"], iteration=st["iteration"], ist=st["ist"],
                            oec=st["oec"])
            try:
                plain.advance_dt(dt, 1)
            except HydroError:
                continue
            r = flux_divergence(net.forward(base).astype(np.float64), scale, dx)
            up = _global(plain.get_state()["conserved"], nb, bs)
            ut = up.copy()
            ut[..., :3] += dt * (r - b)
            truth.set_state(_blocks(ut, nb, bs))
            net.add_sample(base, plain, truth, dt)
            added += 1
    finally:
        for sv in (c, base, plain, truth):
            sv.close()
    return added, (float(np.mean(l1)) if l1 else float("nan")), failed


def train_on_fine(fine_w, coarse_blocks=8, train_orbits=2.0, sample_every=100, max_samples=200,
                  train_steps=4000, lr=1e-3, seed=5, limiter=0.1, train_mask=0.05, pair_steps=1,
                  loop_rounds=0, loop_starts=8, loop_steps=200, loop_every=10, loop_train_steps=2000, burn_in=0,
                  rollout_rounds=0, rollout_steps=200, rollout_every=20, rollout_train_steps=2000, rollout_gain=0.1,
                  rollout_smooth=8, pair_sinks=True, pair_start_orbits=0.0, extra_runs=(), main_pairs=True,
                  main_orbits=None, autotune=True, log=print):
    """Train a closure for the coarse grid (coarse_blocks^2 blocks, 4x coarser
    than the fine workload) on pairs taken every sample_every fine steps along
    the fine run up to train_orbits: each pair compares pair_steps coarse steps
    with the fine run over the same time, from the same down-sampled state, and
    the loss runs over cells with density >= train_mask.  With burn_in > 0 a
    pair instead starts from a coarse run stepped burn_in times beside the
    fine run, and its target is the difference of the two runs' increments
    over the next coarse step (their tendencies at that state).  Then loop_rounds
    rounds of model-in-the-loop pairs (loop_pairs) from loop_starts fine
    states spread over the training window, each round followed by
    loop_train_steps more training steps.  Then rollout_rounds rounds of
    rollout-objective pairs (rollout_pairs: rollout_steps corrected steps from
    each of the loop_starts states, a pair every rollout_every steps), each
    round adding to the samples (the one-step pairs stay) and followed by
    rollout_train_steps training steps; rollout_gain and rollout_smooth damp
    and smooth their targets.  With pair_sinks False the one-step pairs are
    stepped with the sinks off in both grids (from states of the run with
    sinks): a flux correction can only move mass, so a target that holds
    the two grids' different sink removals would teach it to move mass out
    of the sink cells (or into them) instead of correcting the transport;
    the coarse run's own sinks then do the accreting.  Pairs are taken only
    from pair_start_orbits on (the fine run advances through the earlier
    part without them).  extra_runs (fine workloads on the same mesh, with
    their own binary and gas parameters) add their one-step pairs first,
    taken the same way along [pair_start_orbits, train_orbits] of each; with
    main_pairs False the fine workload itself gives none (it is then only
    advanced, to main_orbits if given, else to train_orbits, and a closure
    evaluated on it has never seen its parameters).  The sample capacity is
    shared equally.  autotune: every run on the stage kernel the autotuner
    picks for its grid (else the default tile).  Returns a dict with
    the fine solver at the end of the training window, the trained closure
    (limiter kappa and mask as given), the coarse mesh, hydro, physics, the
    coarse initial state, the trained parameters and output scale, and the
    training figures."""
    from .config import Mesh
    from .solver import Snapshot, Solver, resample_state
    from . import workloads

    fm, hydro, ph = fine_w["mesh"], fine_w["hydro"], fine_w["physics"]
    u0 = workloads.initial_state(fine_w)
    cm = Mesh(coarse_blocks, fm.num_blocks * fm.block_size // (4 * coarse_blocks), fm.domain_radius)
    cu0, _, _ = resample_state(fm, cm, u0)
    fine = Solver(fm, hydro, ph, u0)
    base, truth = (Solver(cm, hydro, ph, cu0) for _ in range(2))
    # the solvers that take the pair steps (the same physics, or without sinks)
    ph_pair = ph
    if not pair_sinks:
        import copy
        ph_pair = copy.deepcopy(ph)
        ph_pair.sink_rate = 0.0
    plain = Solver(cm, hydro, ph_pair, cu0)
    fine_pair = None if pair_sinks else Solver(fm, hydro, ph_pair, u0)
    # every run of the study on the stage kernel the autotuner picks for its
    # grid (the plain fine run is the learned runs' wall-time reference)
    if autotune:
        fine.autotune(reps=3)
        base.autotune(reps=3)
    kf, kc = fine.get_kernel(), base.get_kernel()
    for sv in (truth, plain):
        sv.set_kernel(*kc)
    if fine_pair is not None:
        fine_pair.set_kernel(*kf)
    log(f"stage kernels: fine {kf}, coarse {kc}")
    N = cm.num_blocks * cm.block_size
    dxc = 2.0 * cm.domain_radius / N
    per_round = loop_starts * (-(-loop_steps // loop_every))
    per_rollout = loop_starts * (-(-rollout_steps // rollout_every))
    net = Closure(N, cm.domain_radius, max_samples=max_samples + loop_rounds * per_round +
                  rollout_rounds * per_rollout, seed=seed,
                  init_output_gain=0.1, limiter=limiter, mask_density=train_mask)
    snaps = []
    snap_dt = t_snap = train_orbits * ORBIT / loop_starts if loop_rounds > 0 or rollout_rounds > 0 else None

    def pair(base, plain, truth, fine, fine_pair=fine_pair):
        """Fill base, plain and truth for one training pair; returns its dt."""
        base.resample_from(fine)
        if burn_in > 0:
            # tendency target: the base state is a coarse run stepped burn_in
            # times beside the fine run, and the target is the difference of
            # the two runs' increments over the next step
            for _ in range(burn_in):
                step_beside(base, fine)
            truth.resample_from(fine)
            t0 = truth.get_state()["conserved"]
            plain.resample_from(base)
            dtc = step_beside(plain, fine)
            truth.resample_from(fine)
            t1 = truth.get_state()["conserved"]
            truth.set_state(base.get_state()["conserved"] + (t1 - t0))
            return dtc
        plain.resample_from(fine)
        # pair_steps coarse steps of dtc against 4 pair_steps fine steps of dtc / 4:
        # the target is the mean rate of the difference over that window
        # (without sinks: the fine steps from a copy of the fine state, which
        # itself goes on with its sinks)
        fp = fine
        if fine_pair is not None:
            fine_pair.resample_from(fine)
            fp = fine_pair
        dtc = min(plain.advance(0), 4.0 * fp.advance(0))
        plain.advance_dt(dtc, pair_steps)
        fp.advance_dt(dtc / 4.0, 4 * pair_steps)
        truth.resample_from(fp)
        return dtc * pair_steps

    skipped = 0
    t_train = train_orbits * ORBIT
    if main_orbits is not None and main_pairs:
        raise ValueError("main_orbits applies only when the fine workload gives no pairs")
    t_main = t_train if main_orbits is None else main_orbits * ORBIT
    tp0 = time.perf_counter()
    n_sources = len(extra_runs) + (1 if main_pairs else 0)
    quota = max_samples // max(1, n_sources)
    # one-step pairs from the other runs (their own gas and binary)
    for w in extra_runs:
        if w["mesh"].num_blocks * w["mesh"].block_size != fm.num_blocks * fm.block_size:
            raise ValueError("extra training runs must share the fine mesh")
        ew0 = workloads.initial_state(w)
        ecu0, _, _ = resample_state(w["mesh"], cm, ew0)
        eph_pair = w["physics"]
        if not pair_sinks:
            import copy
            eph_pair = copy.deepcopy(w["physics"])
            eph_pair.sink_rate = 0.0
        ef = Solver(w["mesh"], w["hydro"], w["physics"], ew0)
        eb, et = (Solver(cm, w["hydro"], w["physics"], ecu0) for _ in range(2))
        ep = Solver(cm, w["hydro"], eph_pair, ecu0)
        efp = None if pair_sinks else Solver(w["mesh"], w["hydro"], eph_pair, ew0)
        for sv in (ef,) + ((efp,) if efp is not None else ()):
            sv.set_kernel(*kf)
        for sv in (eb, et, ep):
            sv.set_kernel(*kc)
        start = net.samples()[0]
        log(f"training pairs from a run with e {w['physics'].binary_eccentricity:.3f}, "
            f"q {w['physics'].binary_mass_ratio:.3f}, Mach {w['hydro'].mach_number:.1f}")
        try:
            while ef.time() < t_train and net.samples()[0] - start < quota:
                for _ in range(max(1, sample_every // 10)):
                    ef.advance(10)
                if ef.time() < pair_start_orbits * ORBIT:
                    continue
                try:
                    dtc = pair(eb, ep, et, ef, fine_pair=efp)
                except HydroError as e:
                    skipped += 1
                    log(f"  no pair at t = {ef.time() / ORBIT:.3f} orbits: {e}")
                    continue
                net.add_sample(eb, ep, et, dtc)
        except HydroError as e:
            log(f"  the run stopped at t = {ef.time() / ORBIT:.3f} orbits: {e}")
        finally:
            for sv in (ef, eb, et, ep) + ((efp,) if efp is not None else ()):
                sv.close()
    # training pairs along the fine run
    if main_pairs:
        log(f"training pairs: fine {fm.num_blocks * fm.block_size}^2 to {train_orbits} orbits, coarse {N}^2")
    else:
        log(f"the evaluated run to {t_main / ORBIT:g} orbits, without pairs")
    main_cap = net.samples()[0] + quota if main_pairs else 0
    while fine.time() < t_main:
        for _ in range(max(1, sample_every // 10)):
            fine.advance(10)
        if snap_dt is not None and len(snaps) < loop_starts and fine.time() >= t_snap - snap_dt:
            snaps.append(Snapshot(fine))
            fine.save(snaps[-1])
            t_snap += snap_dt
        if net.samples()[0] >= min(max_samples, main_cap) or fine.time() < pair_start_orbits * ORBIT:
            continue
        try:
            dtc = pair(base, plain, truth, fine)
        except HydroError as e:
            # the coarse scheme fails from this down-sampled state (or the
            # fine run's sub-steps do, leaving it unchanged): no pair
            skipped += 1
            log(f"  no pair at t = {fine.time() / ORBIT:.3f} orbits: {e}")
            continue
        net.add_sample(base, plain, truth, dtc)
        k = net.samples()[0]
        if k % 20 == 0:
            log(f"  {k} samples, t = {fine.time() / ORBIT:.3f} orbits")
    for sv in (base, plain, truth) + ((fine_pair,) if fine_pair is not None else ()):
        sv.close()
    tpairs = time.perf_counter() - tp0
    nsamp, ms = net.samples()
    scale = np.sqrt(ms) * dxc
    net.set_scales(output_scale=scale, loss_weight=1.0 / ms)
    log(f"{nsamp} samples in {tpairs:.1f} s; target rms {np.sqrt(ms)}")

    tt0 = time.perf_counter()
    hist = net.train(train_steps, lr=lr, seed=seed)
    ttrain = time.perf_counter() - tt0
    log(f"trained {train_steps} steps in {ttrain:.1f} s: loss {hist[:50].mean():.4f} -> {hist[-200:].mean():.4f}")
    loop_info, rollout_info = [], []
    if snaps:
        fl = Solver(fm, hydro, ph, u0)
        for rnd in range(loop_rounds):
            tl0 = time.perf_counter()
            added, failed = 0, 0
            for sn in snaps:
                a, f = loop_pairs(net, fl, sn, cm, hydro, ph, cu0, steps=loop_steps, every=loop_every)
                added += a
                failed += f
            h2 = net.train(loop_train_steps, lr=lr, seed=seed + 1 + rnd)
            loop_info.append({"pairs": added, "failed_runs": failed, "loss_last200": float(h2[-200:].mean()),
                              "s": time.perf_counter() - tl0})
            log(f"loop round {rnd}: {added} pairs ({failed} runs failed), loss {h2[-200:].mean():.4f}")
        for rnd in range(rollout_rounds):
            tl0 = time.perf_counter()
            added, failed, l1 = 0, 0, []
            for sn in snaps:
                a, e, f = rollout_pairs(net, fl, sn, cm, hydro, ph, cu0, scale, steps=rollout_steps,
                                        every=rollout_every, gain=rollout_gain, smooth=rollout_smooth)
                added += a
                failed += f
                l1.append(e)
            h2 = net.train(rollout_train_steps, lr=lr, seed=seed + 101 + rnd) if added else np.zeros(1)
            rollout_info.append({"pairs": added, "failed_runs": failed, "rollout_l1_sigma": float(np.nanmean(l1)),
                                 "loss_last200": float(h2[-200:].mean()), "s": time.perf_counter() - tl0})
            log(f"rollout round {rnd}: rollout L1(Sigma) {np.nanmean(l1):.4e} before training, {added} pairs "
                f"({failed} runs failed), loss {h2[-200:].mean():.4f}")
        if rollout_rounds > 0:
            # the rollout loss after the last round
            l1 = []
            for sn in snaps:
                _, e, _ = rollout_pairs(net, fl, sn, cm, hydro, ph, cu0, scale, steps=rollout_steps,
                                        every=rollout_steps, min_fit=rollout_steps + 1)
                l1.append(e)
            rollout_info.append({"rollout_l1_sigma": float(np.nanmean(l1))})
            log(f"rollout L1(Sigma) after training {np.nanmean(l1):.4e}")
        fl.close()
        for sn in snaps:
            sn.close()
    return {"fine": fine, "net": net, "coarse_mesh": cm, "hydro": hydro, "physics": ph, "cu0": cu0,
            "kernel_fine": kf, "kernel_coarse": kc,
            "params": net.params(), "scale": scale,
            "info": {"fine_grid": fm.num_blocks * fm.block_size, "coarse_grid": N, "samples": nsamp,
                     "train_steps": train_steps, "loss_first50": float(hist[:50].mean()),
                     "loss_last200": float(hist[-200:].mean()), "pairs_s": tpairs, "train_s": ttrain,
                     "skipped_pairs": skipped, "loop_rounds": loop_info, "rollout_rounds": rollout_info}}


def _perturbed(solver, rel, seed):
    """The solver's committed state with every conserved value multiplied by
    (1 + rel n), n uniform in [-1, 1] from a seeded generator (time,
    iteration, block sums and torque totals kept)."""
    st = solver.get_state()
    rng = np.random.default_rng(seed)
    u = st["conserved"] * (1.0 + rel * rng.uniform(-1.0, 1.0, st["conserved"].shape))
    solver.set_state(u, time=st["time"], iteration=st["iteration"], ist=st["ist"], oec=st["oec"],
                     torque=solver.torque())


def mdot_spread(fine, snap, coarse_mesh, hydro, physics, cu0, t_end, runs, rel=1e-10, kernel=None, log=print):
    """Time-averaged accretion rates of `runs` plain fine runs (the fine
    solver restored from snap, its state perturbed by _perturbed with seeds
    1..runs) and of plain coarse runs from each perturbed state
    down-sampled, over [snap's time, t_end], stepped as learned_vs_fine steps
    its plain runs.  Returns the rates and their relative spread (max - min
    over the mean) for both grids; kernel: the coarse runs' stage kernel."""
    from .solver import Solver
    fr, cr = [], []
    coarse = Solver(coarse_mesh, hydro, physics, cu0)
    if kernel is not None:
        coarse.set_kernel(*kernel)
    try:
        for k in range(1, runs + 1):
            fine.restore(snap)
            _perturbed(fine, rel, k)
            coarse.resample_from(fine)
            for rates, s in ((cr, coarse), (fr, fine)):
                m0, ta = accreted(s), s.time()
                run_to(lambda v: 10 * v.advance(10), s, t_end, fold=10)
                rates.append((accreted(s) - m0) / (s.time() - ta))
            log(f"spread run {k}: fine mdot {fr[-1]:.6e}, coarse mdot {cr[-1]:.6e}")
    finally:
        coarse.close()
    def spread(r):
        m = float(np.mean(r))
        return float((max(r) - min(r)) / m) if m != 0.0 else 0.0
    return {"rel_perturbation": rel, "fine_mdot": fr, "coarse_mdot": cr, "fine_rel_spread": spread(fr),
            "coarse_rel_spread": spread(cr)}


def _sink_rates(mesh, physics, t):
    """Physics::sink_kernel (src/physics.rs:388-397) of both bodies at time t
    at every cell centre of mesh, [N][N] (i-major)."""
    from .solver import orbital_state_from_time
    N = mesh.num_blocks * mesh.block_size
    dx = 2.0 * mesh.domain_radius / N
    x = -mesh.domain_radius + dx * (np.arange(N) + 0.5)
    X, Y = np.meshgrid(x, x, indexing="ij")
    st = orbital_state_from_time(physics, t)
    s2 = physics.sink_radius ** 2
    k = np.zeros_like(X)
    for b in range(2):
        r2 = (X - st[b, 1]) ** 2 + (Y - st[b, 2]) ** 2
        k += physics.sink_rate * np.where(r2 < s2, (1.0 - r2 / s2) ** 2, 0.0)
    return k, dx


def sink_quadrature(fine_w, coarse_blocks=8, t0_orbits=2.0, t1_orbits=22.0, every_orbits=0.25, log=print):
    """How the coarse grid's sinks sample a density field, apart from its
    transport: along the plain fine run over [t0_orbits, t1_orbits], at the
    end of the first 10-step call past every `every_orbits`, the
    mass-removal rate of the sinks (sum over cells of sink rate x Sigma x
    cell area) on the fine grid, and on the coarse grid (4x coarser) with the
    fine Sigma down-sampled.  (The samples fall where the calls end, at
    orbital phases that drift: sampling at exact fractions of an orbit would
    put a circular binary at the same few places on the grid every time.)  A
    closure that reproduced the down-sampled fine state exactly would accrete
    at the coarse rate; its ratio to the fine rate is the bias left to it.
    Returns the per-state rates, their ratios and the ratio of the window
    means."""
    from .config import Mesh
    from .solver import Solver, resample_state
    from . import workloads
    fm, hydro, ph = fine_w["mesh"], fine_w["hydro"], fine_w["physics"]
    cm = Mesh(coarse_blocks, fm.num_blocks * fm.block_size // (4 * coarse_blocks), fm.domain_radius)
    fine = Solver(fm, hydro, ph, workloads.initial_state(fine_w))
    rows = []
    try:
        t_next = t0_orbits * ORBIT
        while t_next <= t1_orbits * ORBIT * (1.0 + 1e-12):
            fine.advance(10)
            if fine.time() < t_next:
                continue
            t = fine.time()
            u = fine.get_state()["conserved"]
            uc, _, _ = resample_state(fm, cm, u)
            kf, dxf = _sink_rates(fm, ph, t)
            kc, dxc = _sink_rates(cm, ph, t)
            mf = float((kf * _global(u, fm.num_blocks, fm.block_size)[..., 0]).sum() * dxf * dxf)
            mc = float((kc * _global(uc, cm.num_blocks, cm.block_size)[..., 0]).sum() * dxc * dxc)
            rows.append([t / ORBIT, mf, mc])
            if len(rows) % 8 == 0:
                log(f"  t = {t / ORBIT:.2f} orbits: fine {mf:.4e}, coarse sampling {mc:.4e}")
            t_next += every_orbits * ORBIT
    finally:
        fine.close()
    a = np.array(rows)
    r = a[:, 2] / np.where(a[:, 1] > 0, a[:, 1], np.nan)
    return {"orbits": [t0_orbits, t1_orbits], "every_orbits": every_orbits, "states": len(rows),
            "fine_rate": a[:, 1].tolist(), "coarse_sampling_rate": a[:, 2].tolist(), "time_orbits": a[:, 0].tolist(),
            "ratio_min": float(np.nanmin(r)), "ratio_max": float(np.nanmax(r)),
            "ratio_of_means": float(a[:, 2].mean() / a[:, 1].mean())}


def variant_fields(v):
    """(kappa, damping, mask density, components, interval) of a learned-run
    variant given as 3 to 5 numbers: components default 0 (all), interval 1
    (a correction every step)."""
    v = tuple(float(x) for x in v)
    if not 3 <= len(v) <= 5:
        raise ValueError(f"a variant has 3 to 5 fields, not {len(v)}")
    return v[0], v[1], v[2], int(v[3]) if len(v) > 3 else 0, int(v[4]) if len(v) > 4 else 1


def learned_vs_fine(fine_w, coarse_blocks=8, train_orbits=2.0, eval_orbits=1.0, sample_every=100,
                    max_samples=200, train_steps=4000, lr=1e-3, seed=5, limiter=0.1, train_mask=0.05,
                    pair_steps=1, variants=((0.1, 1.0, 0.05),), heldout_pairs=16, loop_rounds=0, loop_starts=8,
                    loop_steps=200, loop_every=10, loop_train_steps=2000, burn_in=0, learned_fold=10,
                    rollout_rounds=0, rollout_steps=200, rollout_every=20, rollout_train_steps=2000,
                    rollout_gain=0.1, rollout_smooth=8, pair_sinks=True, pair_start_orbits=0.0, spread_runs=0,
                    train_runs=(), eval_start_orbits=None, net_seeds=(), log=print):
    """Train a closure on the fine workload (train_on_fine), measure it on the
    fine states that follow the training window (heldout_one_step, from a
    device snapshot of the fine run), then run plain fine, plain coarse and
    learned coarse runs over the held-out window
    [train_orbits, train_orbits + eval_orbits] from the same fine state; one
    learned run per (limiter kappa, damping, mask density[, components])
    variant, the damping scaling the trained output, the fourth entry
    keeping only the mass-flux correction (1) or only the momentum-flux
    corrections (2) and the fifth correcting every that many steps
    (Closure.set_interval).  Every run ends exactly at the
    window's end.  Reported per run: wall time per orbit, the time-averaged
    accretion rate and torque on the binary over the window, and for the
    coarse runs the L1 error of Sigma at the window's end against the
    down-sampled fine state, with the relative errors of the rates against
    the fine run.  With spread_runs > 0, that many more plain fine and plain
    coarse runs over the same window start from the window's starting state
    with a seeded relative perturbation of 1e-10 in every conserved field:
    their time-averaged accretion rates (`mdot_spread`) show how far a run's
    own rate moves when only the rounding of its start differs (the flow is
    chaotic over many orbits), the scale against which an error of a few
    percent in a 20-orbit mean is to be read.  With train_runs (fine
    workloads on the same mesh) the closure is trained on their one-step
    pairs only, and fine_w, whose parameters it has then never seen, serves
    for the held-out states and the window, which then starts at
    eval_start_orbits if given (the training runs may run longer).
    net_seeds: more nets trained afresh from those seeds' initial weights on
    the same one-step pairs (no loop or rollout rounds), each run over the
    window in every variant (names suffixed _s<seed>); the plain runs are
    shared."""
    from .solver import Snapshot, Solver, resample_state

    if net_seeds and (loop_rounds or rollout_rounds):
        raise ValueError("net_seeds retrain the one-step pairs only (no loop or rollout rounds)")
    if train_runs and (loop_rounds or rollout_rounds):
        raise ValueError("loop and rollout rounds start from states of the evaluated run: with train_runs "
                         "they would train on the run they are evaluated on")
    t = train_on_fine(fine_w, coarse_blocks=coarse_blocks, train_orbits=train_orbits, sample_every=sample_every,
                      max_samples=max_samples, train_steps=train_steps, lr=lr, seed=seed, limiter=limiter,
                      train_mask=train_mask, pair_steps=pair_steps, loop_rounds=loop_rounds,
                      loop_starts=loop_starts, loop_steps=loop_steps, loop_every=loop_every,
                      loop_train_steps=loop_train_steps, burn_in=burn_in, rollout_rounds=rollout_rounds,
                      rollout_steps=rollout_steps, rollout_every=rollout_every,
                      rollout_train_steps=rollout_train_steps, rollout_gain=rollout_gain,
                      rollout_smooth=rollout_smooth, pair_sinks=pair_sinks, pair_start_orbits=pair_start_orbits,
                      extra_runs=train_runs, main_pairs=not train_runs,
                      main_orbits=eval_start_orbits if train_runs else None, log=log)
    fine, net, cm, hydro, ph, cu0 = t["fine"], t["net"], t["coarse_mesh"], t["hydro"], t["physics"], t["cu0"]
    trainer = net
    fm = fine_w["mesh"]
    N = cm.num_blocks * cm.block_size
    params, trained_scale = t["params"], t["scale"]
    out = dict(t["info"])
    if heldout_pairs > 0:
        snap = Snapshot(fine)
        fine.save(snap)
        out["heldout_one_step"] = heldout_one_step(fine, net, cm, hydro, ph, cu0, pairs=heldout_pairs,
                                                   burn_in=burn_in)
        fine.restore(snap)
        snap.close()
        log(f"held-out one step: L1 ratio learned / plain {out['heldout_one_step']['ratio']}")

    # held-out window from the fine state at the end of training
    t_start = fine.time()
    t_end = t_start + eval_orbits * ORBIT
    kc = t["kernel_coarse"]
    start = Solver(cm, hydro, ph, cu0)
    start.resample_from(fine)
    coarse = Solver(cm, hydro, ph, cu0)
    coarse.set_kernel(*kc)
    coarse.resample_from(fine)
    out["stage_kernels"] = {"fine": list(t["kernel_fine"]), "coarse": list(kc)}
    out["window_orbits"] = [t_start / ORBIT, t_end / ORBIT]
    snap0 = None
    if spread_runs > 0:
        snap0 = Snapshot(fine)
        fine.save(snap0)
    # plain runs step as the reference's driver does (fold = 10 steps per
    # advance at one dt); the learned runs correct every step, learned_fold
    # corrected steps per call (one replayed graph)
    runs = [("fine", fine, lambda s: 10 * s.advance(10), None),
            ("coarse", coarse, lambda s: 10 * s.advance(10), None)]
    seed_params = [(seed, "", params)]
    for sd in net_seeds:
        # a fresh net's initial weights for this seed, trained on the same pairs
        init = Closure(N, cm.domain_radius, seed=sd, init_output_gain=0.1)
        p0 = init.params()
        init.close()
        trainer.set_params(p0)
        tt0 = time.perf_counter()
        h = trainer.train(train_steps, lr=lr, seed=sd)
        log(f"net seed {sd}: trained {train_steps} steps in {time.perf_counter() - tt0:.1f} s: "
            f"loss {h[:50].mean():.4f} -> {h[-200:].mean():.4f}")
        out[f"loss_last200_s{sd}"] = float(h[-200:].mean())
        seed_params.append((sd, f"_s{sd}", trainer.params()))
    for sd, suffix, prm in seed_params:
        for v in variants:
            kappa, damping, mask, mass_only, every = variant_fields(v)
            name = (f"learned_k{kappa:g}_d{damping:g}_m{mask:g}" + {0: "", 1: "_mass", 2: "_mom"}[mass_only]
                    + (f"_i{every}" if every > 1 else "") + suffix)
            runs.append((name, None, lambda s: learned_fold * net.advance(s, learned_fold),
                         (kappa, damping, mask, mass_only, prm, every)))
    sig_fine = None
    for name, solver, step, var in runs:
        if var is not None:
            if net is not trainer:
                net.close()
            net = Closure(N, cm.domain_radius, limiter=var[0], mask_density=var[2])
            net.set_params(var[4])
            # the corrected flux components: all, the mass flux only (1) or
            # the momentum fluxes only (2)
            comp = {0: np.ones(3), 1: np.array([1.0, 0.0, 0.0]), 2: np.array([0.0, 1.0, 1.0])}[int(var[3])]
            net.set_scales(output_scale=trained_scale * var[1] * comp)
            net.set_interval(var[5])
            solver = Solver(cm, hydro, ph, cu0)
            solver.set_kernel(*kc)
            solver.resample_from(start)
        m0, ta, tq0 = accreted(solver), solver.time(), solver.torque().sum()
        try:
            n, wall = run_to(step, solver, t_end, log=lambda m: log(f"{name}:{m}"), every=1000,
                             fold=10 if var is None else learned_fold)
        except (HydroError, Stalled) as e:
            log(f"{name}: failed: {e}")
            out[name] = {"failed": str(e)}
            continue
        tb = solver.time()
        mdot = (accreted(solver) - m0) / (tb - ta)
        torque = (solver.torque().sum() - tq0) / (tb - ta)
        u = solver.get_state()["conserved"]
        out[name] = {"advance_calls": n, "wall_s": wall, "orbits": (tb - ta) / ORBIT,
                     "wall_per_orbit_s": wall / ((tb - ta) / ORBIT), "mdot": mdot, "torque": torque}
        if name == "fine":
            sig_fine = resample_state(fm, cm, u)[0][..., 0]
        elif sig_fine is not None:
            out[name]["sigma_l1_rel"] = float(np.abs(u[..., 0] - sig_fine).sum() / np.abs(sig_fine).sum())
        log(f"{name}: {n} calls, {wall:.2f} s, mdot {mdot:.6e}, torque {torque:.6e}")
    if snap0 is not None:
        out["mdot_spread"] = mdot_spread(fine, snap0, cm, hydro, ph, cu0, t_end, spread_runs, kernel=kc, log=log)
        snap0.close()
    if "mdot" in out.get("fine", {}):
        ref, reft = out["fine"]["mdot"], out["fine"]["torque"]
        for name in [r[0] for r in runs[1:]]:
            if "mdot" in out.get(name, {}):
                out[name]["mdot_rel_err"] = abs(out[name]["mdot"] - ref) / abs(ref)
                out[name]["torque_rel_err"] = abs(out[name]["torque"] - reft) / abs(reft)
                out[name]["speedup_vs_fine"] = out["fine"]["wall_per_orbit_s"] / out[name]["wall_per_orbit_s"]
    if net is not trainer:
        net.close()
    trainer.close()
    return out
