"""Benchmark: FP64 zone-updates/s of the per-block RK update on one MI355X.

Workload (default euler2048): gamma-law gas on 2048^2 cells (32x32 blocks of
64^2), RK3, alpha viscosity, eccentric (e = 0.5) unequal (q = 0.5) binary,
beta cooling; seeded cavity initial state, resident in HBM before timing.
A step is one full RK step of the whole mesh (one zone-update per cell, as
the reference's Mzps counts it).  Steps run through the C ABI as
cdc_solver_advance(fold) calls; every call returns after its work is done
(each call replays the solver's captured launch graph).

Also reported: the roofline of the fused stage kernel (algorithmic HBM bytes
per launch over its average duration, timed with HIP events bound to each
stage dispatch in a second pass right after the timed region) and the CPU
oracle on the same workload for a bounded sample, timed on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "app-binary2d_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_step(nv, rk):
    """HBM bytes per cell per full RK step: stage 1 reads U and U_init and
    writes U*; later stages also read the step base."""
    return 8 * nv * (3 + 4 * (rk - 1))


def kernel_name(tile):
    """The stage kernel a tile shape selects: 64-column strips are the
    column-marching kernel's."""
    return "march_kernel_batch" if tile.endswith("x64") else "stage_kernel_batch"


def committed_profile(workload, tile):
    """The committed PMC summary of the stage kernel (HBM traffic from
    calibrated FETCH_SIZE / WRITE_SIZE passes, VALU instructions per wave)
    when it covers this workload and tile; {} otherwise."""
    stem = kernel_name(tile).replace("_batch", "")
    path = os.path.join(ROOT, "profiles", f"{stem}_{workload}_{tile}_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return {}
    if t.get("workload") != workload or t.get("tile") != tile:
        return {}
    return t


def valu_issue(prof, avg_launch_ms):
    """The launch's FP64-VALU issue floor (instructions x 4 cycles per wave64
    instruction over all SIMDs) and the fraction of it achieved."""
    if not prof or "valu_insts_per_wave" not in prof:
        return None
    cycles = prof["waves_per_launch"] * prof["valu_insts_per_wave"] * prof["valu_cycles_per_wave_instruction"]
    floor_ms = cycles / prof["simds"] / (prof["clock_ghz"] * 1e9) * 1e3
    return {"floor_ms": floor_ms, "frac": floor_ms / avg_launch_ms, "source": prof.get("pmc_source")}


def host_cpus():
    """(CPUs of the host, CPUs this process may run on, physical cores of the
    host from /proc/cpuinfo or None)."""
    total = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = total
    phys = set()
    try:
        with open("/proc/cpuinfo") as f:
            pid = cid = None
            for line in f:
                if line.startswith("physical id"):
                    pid = line.split(":", 1)[1].strip()
                elif line.startswith("core id"):
                    cid = line.split(":", 1)[1].strip()
                elif not line.strip():
                    if pid is not None and cid is not None:
                        phys.add((pid, cid))
                    pid = cid = None
    except OSError:
        pass
    return total, allowed, (len(phys) or None)


def default_cpu_threads():
    """Every CPU this process may run on, capped by OMP_NUM_THREADS when the
    environment sets it (a shared host's CPU share for this GPU)."""
    _, allowed, _ = host_cpus()
    try:
        cap = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        cap = 0
    return min(allowed, cap) if cap > 0 else allowed


def cpu_baseline(w, u0, seconds, threads):
    from oracle import oracle as O
    O.set_threads(threads)
    m, h, p = w["mesh"].to_c(), w["hydro"].to_c(), w["physics"].to_c()
    o = O.Oracle(m, h, p, u0, w["hydro"].nv)
    o.advance(1)  # first touch of the work buffers
    n = 0
    t0 = time.perf_counter()
    while True:
        o.advance(1)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    o.close()
    zones = w["mesh"].total_zones()
    cpu = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    total, allowed, phys = host_cpus()
    used = O.get_threads()
    return {"value": n * zones / el, "unit": "zone-updates/s", "cores": used, "kind": "port",
            "sample": f"{n} RK steps of the same workload ({el:.1f} s), C oracle, OpenMP over blocks, "
                      f"{used} threads on {cpu}",
            "mzps": n * zones / el * 1e-6,
            "host": {"cpu": cpu, "logical_cpus": total, "physical_cores": phys, "cpus_allowed": allowed,
                     "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}}


def run_ensemble(args):
    """The 64-run parameter study, every run from t = 0 to half an orbit,
    one ensemble call per step of all unfinished runs."""
    import math
    import numpy as np
    from cdc_amd import Ensemble, Solver, workloads
    members = workloads.ensemble_members()
    inits = [workloads.initial_state(m) for m in members]
    solvers = [Solver(m["mesh"], m["hydro"], m["physics"], u) for m, u in zip(members, inits)]
    # one stage kernel for all members: the given tile, or the autotuner's
    # pick on the largest member (most of the work)
    if args.tile:
        shape = tuple(int(v) for v in args.tile.lower().split("x"))
        for g in solvers:
            g.set_tile(*shape)
    elif not args.no_autotune:
        # candidates: the autotuner's pick on the largest member (most of the
        # work) and, when every block size is a multiple of 64, the
        # column-marching strips; each is timed on the whole ensemble
        big = max(solvers, key=lambda g: g.mesh.total_zones())
        big.autotune(reps=3)
        cands = [big.get_kernel()]
        if all(g.mesh.block_size % 64 == 0 for g in solvers):
            cands += [k for k in ((32, 64, 64), (16, 64, 64), (8, 64, 64)) if k not in cands]
        best = None
        for kern in cands:
            for g in solvers:
                g.set_kernel(*kern)
            with Ensemble(solvers) as trial:
                trial.advance()
                t0 = time.perf_counter()
                for _ in range(3):
                    trial.advance()
                el = time.perf_counter() - t0
            if best is None or el < best[0]:
                best = (el, kern)
        for g, u in zip(solvers, inits):
            g.set_kernel(*best[1])
            g.set_state(u)
    ens = Ensemble(solvers)
    for _ in range(args.warmup):
        ens.advance()
    for g, u in zip(solvers, inits):
        g.set_state(u)
    target = 0.5 * 2.0 * math.pi
    n = len(solvers)
    times = np.zeros(n)
    calls = 0
    z0 = ens.stats()[1]
    t0 = time.perf_counter()
    while True:
        active = (times < target).astype(np.int32)
        if not active.any():
            break
        dt, codes, errs = ens.advance(active)
        if np.any(codes != 0):
            bad = int(np.nonzero(codes)[0][0])
            raise RuntimeError(f"ensemble member {bad} failed: {errs[bad]}")
        times = np.where(active == 1, times + dt, times)
        calls += 1
    wall = time.perf_counter() - t0
    zones = ens.stats()[1] - z0
    exact = np.array([g.time() for g in solvers])
    assert np.all(exact >= target * (1 - 1e-12)), "a run stopped short of half an orbit"
    sizes = sorted({m["mesh"].num_blocks * m["mesh"].block_size for m in members})
    line = {
        "metric": "f64 zone-updates/s, 64-run ensemble to half an orbit",
        "value": zones / wall,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": calls,
        "warmup": args.warmup,
        "ms_per_step": wall / calls * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded cavity disks, parameters drawn with seed 4)",
        "config": {"workload": "ensemble64", "runs": n, "grids": [f"{s}^2" for s in sizes],
                   "hydro": "isothermal", "rk": "RK2", "tile": "%dx%d" % ens.get_tile(),
                   "zone_updates": zones, "wall_s": wall},
    }
    ens.close()
    for g in solvers:
        g.close()
    print(json.dumps(line), flush=True)


def run_sink_quadrature(args):
    """The coarse grid's sink sampling of the fine run's density (see
    closure.sink_quadrature), as one JSON line."""
    from cdc_amd import workloads
    from cdc_amd.closure import sink_quadrature
    w = workloads.get("iso2048")
    log = lambda m: print(m, file=sys.stderr, flush=True)
    r = sink_quadrature(w, t0_orbits=args.train_orbits, t1_orbits=args.train_orbits + args.eval_orbits, log=log)
    print(json.dumps({"metric": "coarse (512^2) sink sampling of the down-sampled 2048^2 density / 2048^2 sinks",
                      "value": r["ratio_of_means"], "unit": "ratio of window means", "n_gpus": 1,
                      "data": "synthetic (seeded cavity disk, seed %d)" % w["seed"],
                      "config": {"workload": "sinkq", "fine": "iso2048", **r}}), flush=True)


def run_closure_step(args):
    """Wall time of corrected and plain steps on the learned closure's 512^2
    isothermal coarse grid (the 2048^2 initial state down-sampled 4x), the
    net with seeded random weights: --steps steps of each after --warmup,
    --learned-fold corrected steps per call and --fold plain steps per call,
    as one JSON line."""
    import numpy as np
    from cdc_amd import Solver, workloads
    from cdc_amd.closure import Closure, num_params
    from cdc_amd.config import Mesh
    from cdc_amd.solver import resample_state
    w = workloads.get("iso2048")
    fm = w["mesh"]
    cm = Mesh(8, 64, fm.domain_radius)
    cu0, _, _ = resample_state(fm, cm, workloads.initial_state(w))
    n = cm.num_blocks * cm.block_size
    net = Closure(n, cm.domain_radius, limiter=0.02, mask_density=0.05)
    rng = np.random.default_rng(args.closure_seed)
    net.set_params((0.05 * rng.standard_normal(num_params())).astype(np.float32))
    # small output scale: the corrections act, the flow stays the plain one's
    net.set_scales(output_scale=[1e-6, 1e-6, 1e-6])
    net.set_interval(args.correction_interval)
    lf, pf = max(1, args.learned_fold), max(1, args.fold)

    def timed(call, per_call):
        for _ in range(max(1, args.warmup // per_call)):
            call()
        k = max(1, args.steps // per_call)
        t0 = time.perf_counter()
        for _ in range(k):
            call()
        return (time.perf_counter() - t0) / (k * per_call), k * per_call

    s = Solver(cm, w["hydro"], w["physics"], cu0)
    p = Solver(cm, w["hydro"], w["physics"], cu0)
    # the stage kernel: the given tile, else the autotuner's pick, both runs
    if args.tile:
        s.set_tile(*(int(v) for v in args.tile.lower().split("x")))
    elif not args.no_autotune:
        s.autotune(reps=3)
    p.set_kernel(*s.get_kernel())
    t_corr, n_corr = timed(lambda: net.advance(s, lf), lf)
    t_plain, n_plain = timed(lambda: p.advance(pf), pf)
    tile = "%dx%d" % p.get_tile()
    net.close(); s.close(); p.close()
    print(json.dumps({"metric": "corrected 512\u00b2 step wall time", "value": t_corr * 1e6, "unit": "us",
                      "n_gpus": 1, "steps": n_corr, "warmup": args.warmup, "higher_is_better": False,
                      "dtype": "f64 (net: bf16 operands, fp32 accumulation)",
                      "data": "synthetic (seeded cavity disk, seed %d, down-sampled 4x)" % w["seed"],
                      "config": {"workload": "closure512", "grid": "512^2", "blocks": "8x8 of 64^2",
                                 "hydro": "isothermal", "rk": w["physics"].rk_order, "learned_fold": lf,
                                 "correction_interval": args.correction_interval,
                                 "plain_fold": pf, "plain_us_per_step": t_plain * 1e6,
                                 "tile": tile,
                                 "plain_steps": n_plain, "net_seed": args.closure_seed}}), flush=True)


def run_learned(args):
    """Learned 512^2 closure vs plain 2048^2 (and plain 512^2), isothermal:
    wall time per orbit, accretion rate, torque and Sigma error over a
    held-out window, and the held-out one-step gain of the closure."""
    from cdc_amd import workloads
    from cdc_amd.closure import learned_vs_fine
    w = workloads.get("iso2048")
    log = lambda m: print(m, file=sys.stderr, flush=True)
    variants = [tuple(float(x) for x in v.split(":")) for v in args.variants.split(",")]
    if args.rollout_rounds is None:
        args.rollout_rounds = 0 if args.train_runs else 3
    r = learned_vs_fine(w, train_orbits=args.train_orbits, eval_orbits=args.eval_orbits,
                        sample_every=args.sample_every, max_samples=args.max_samples,
                        train_steps=args.train_steps, pair_steps=args.pair_steps, variants=variants,
                        loop_rounds=args.loop_rounds, loop_steps=args.loop_steps, loop_every=args.loop_every,
                        burn_in=args.burn_in, learned_fold=args.learned_fold, rollout_rounds=args.rollout_rounds,
                        rollout_steps=args.rollout_steps, rollout_every=args.rollout_every,
                        rollout_train_steps=args.rollout_train_steps, rollout_gain=args.rollout_gain,
                        rollout_smooth=args.rollout_smooth, pair_sinks=not args.pair_no_sinks,
                        seed=args.closure_seed, spread_runs=args.mdot_spread,
                        pair_start_orbits=args.pair_start_orbits,
                        train_runs=workloads.training_runs(args.train_runs) if args.train_runs else (),
                        eval_start_orbits=args.eval_start_orbits,
                        net_seeds=[int(x) for x in args.net_seeds.split(",")] if args.net_seeds else (), log=log)
    # the headline: the fastest learned variant within 2% accretion-rate error
    ok = [k for k, v in r.items() if k.startswith("learned") and isinstance(v, dict)
          and v.get("mdot_rel_err", 1.0) <= 0.02]
    best = max(ok, key=lambda k: r[k]["speedup_vs_fine"]) if ok else None
    line = {
        "metric": "learned 512\u00b2 vs plain 2048\u00b2 wall/orbit @\u22642% \u1e40 err",
        "value": r[best]["speedup_vs_fine"] if best else None,
        "unit": "x (plain 2048\u00b2 wall per orbit / learned 512\u00b2 wall per orbit)",
        "n_gpus": 1, "steps": r[best]["advance_calls"] if best else None, "warmup": 0,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (net: bf16 operands, fp32 accumulation)",
        "data": "synthetic (seeded cavity disk, seed %d); training pairs from %s" % (
            w["seed"], "%d runs with drawn parameters" % args.train_runs if args.train_runs else "the same run"),
        "config": {"workload": "learned512", "fine": "iso2048 (isothermal, 32x32 blocks of 64^2, RK2)",
                   **r},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--fold", type=int, default=10)
    ap.add_argument("--timing-steps", type=int, default=30,
                    help="steps of the second pass that times each stage launch with HIP events")
    ap.add_argument("--workload", default="euler2048")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=None,
                    help="oracle threads (default: every CPU this process may run on, capped by OMP_NUM_THREADS)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tile", default=None, help="TXxTY stage-kernel tile (skips the autotuner)")
    ap.add_argument("--train-orbits", type=float, default=2.0)
    ap.add_argument("--rollout-rounds", type=int, default=None,
                    help="rounds of rollout-objective training pairs after the one-step training "
                         "(default 3; 0 with --train-runs, whose rounds would start from the evaluated run)")
    ap.add_argument("--rollout-steps", type=int, default=200, help="corrected coarse steps per rollout")
    ap.add_argument("--rollout-every", type=int, default=20, help="rollout steps between training pairs")
    ap.add_argument("--rollout-train-steps", type=int, default=2000, help="training steps per rollout round")
    ap.add_argument("--rollout-gain", type=float, default=0.1, help="step on the fitted rollout error growth")
    ap.add_argument("--rollout-smooth", type=int, default=8, help="box filter (cells) of that growth")
    ap.add_argument("--learned-fold", type=int, default=10,
                    help="corrected coarse steps per call of the learned runs")
    ap.add_argument("--eval-orbits", type=float, default=2.0)
    ap.add_argument("--closure-seed", type=int, default=5, help="seed of the net's initial weights and batches")
    ap.add_argument("--correction-interval", type=int, default=1,
                    help="closure512: correct every that many steps")
    ap.add_argument("--pair-start-orbits", type=float, default=0.0, help="training pairs only from this time on")
    ap.add_argument("--train-runs", type=int, default=0,
                    help="train on this many 2048^2 runs with drawn parameters instead of the evaluated run")
    ap.add_argument("--eval-start-orbits", type=float, default=None,
                    help="with --train-runs: the evaluated run's window starts here (default: --train-orbits)")
    ap.add_argument("--net-seeds", default="",
                    help="more net seeds trained on the same one-step pairs, comma-separated")
    ap.add_argument("--mdot-spread", type=int, default=0,
                    help="plain fine and coarse runs from perturbed copies of the window's start")
    ap.add_argument("--pair-no-sinks", action="store_true",
                    help="one-step training pairs stepped with the sinks off in both grids")
    ap.add_argument("--train-steps", type=int, default=4000)
    ap.add_argument("--pair-steps", type=int, default=1, help="coarse steps per training pair")
    ap.add_argument("--sample-every", type=int, default=100, help="fine steps between training pairs")
    ap.add_argument("--max-samples", type=int, default=200, help="training pairs kept")
    ap.add_argument("--loop-rounds", type=int, default=0, help="rounds of model-in-the-loop training pairs")
    ap.add_argument("--loop-steps", type=int, default=200, help="corrected coarse steps per loop run")
    ap.add_argument("--loop-every", type=int, default=10, help="coarse steps between loop training pairs")
    ap.add_argument("--burn-in", type=int, default=0,
                    help="coarse steps beside the fine run before each pair (tendency targets)")
    ap.add_argument("--variants", default="0.02:0.25:0.05,0.02:0.5:0.05,0.05:0.5:0.05",
                    help="learned runs as kappa:damping:mask_density[:components[:interval]],...; components "
                         "1: the mass flux only, 2: the momentum fluxes only; interval: correct every that many "
                         "steps")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only; --gpus {args.gpus} is not supported (the mesh would "
              f"shard as 2-D sub-domains with one halo exchange per RK stage, not implemented)",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "ensemble64":
        run_ensemble(args)
        return
    if args.workload == "learned512":
        run_learned(args)
        return
    if args.workload == "sinkq":
        run_sink_quadrature(args)
        return
    if args.workload == "closure512":
        run_closure_step(args)
        return
    from cdc_amd import HydroError, Solver, workloads
    w = workloads.get(args.workload)
    u0 = workloads.initial_state(w)
    mesh, hydro, ph = w["mesh"], w["hydro"], w["physics"]
    zones = mesh.total_zones()
    fold = max(1, min(args.fold, args.steps))
    steps = (args.steps // fold) * fold
    g = Solver(mesh, hydro, ph, u0)
    if args.tile:
        g.set_tile(*(int(v) for v in args.tile.lower().split("x")))
    elif not args.no_autotune:
        g.autotune(reps=3)
    # warmup with the chosen tile: also captures the launch graph that the
    # timed advances replay
    done = 0
    while done < args.warmup:
        g.advance(min(fold, args.warmup - done))
        done += min(fold, args.warmup - done)

    timing_steps = max(fold, (args.timing_steps // fold) * fold)
    if args.workload == "euler2048" and args.warmup + steps + timing_steps > 400:
        print("bench.py: the 2048^2 alpha-viscosity workload loses stability after about 470 steps "
              "(the CFL step has no viscous bound); keep warmup + steps + timing steps below 400",
              file=sys.stderr)

    def run(n):
        try:
            for _ in range(n // fold):
                g.advance(fold)
        except HydroError as e:
            print(json.dumps({"metric": "f64 zone-updates/s", "value": None, "error": str(e),
                              "config": {"workload": args.workload}}), flush=True)
            sys.exit(1)

    # timed region: each advance replays the solver's captured launch graph
    t0 = time.perf_counter()
    run(steps)
    t1 = time.perf_counter()
    wall = t1 - t0
    # stage-kernel duration: HIP events bound to each stage dispatch, over a
    # second timed pass launched directly (events recorded inside a captured
    # graph cannot be timed on ROCm); the kernel is the same machine code
    g.set_timing(True)
    run(timing_steps)
    launches, kernel_ms, cells = g.get_timing()
    prof = committed_profile(args.workload, "%dx%d" % g.get_tile())
    rk = ph.rk_substeps()
    bytes_total = algorithmic_bytes_per_step(hydro.nv, rk) * zones * timing_steps
    achieved = bytes_total / (kernel_ms * 1e-3) / 1e9
    line = {
        "metric": "f64 zone-updates/s @2048²; learned 512² vs plain 2048² wall/orbit @≤2% Ṁ err",
        "value": steps * zones / wall,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": wall / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded cavity disk, seed %d)" % w["seed"],
        "config": {"workload": args.workload, "grid": f"{mesh.num_blocks * mesh.block_size}^2",
                   "blocks": f"{mesh.num_blocks}x{mesh.num_blocks} of {mesh.block_size}^2",
                   "hydro": type(hydro).__name__.lower(), "rk": ph.rk_order, "fold": fold,
                   "tile": "%dx%d" % g.get_tile()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": prof.get("traffic_bytes_per_launch"),
                     "kernel": kernel_name("%dx%d" % g.get_tile()), "launches": launches,
                     "avg_launch_ms": kernel_ms / max(launches, 1),
                     "bytes_per_launch": bytes_total / max(launches, 1),
                     "valu_issue": valu_issue(prof, kernel_ms / max(launches, 1)),
                     "timing": "HIP events bound to each stage dispatch, %d steps after the timed region"
                               % timing_steps,
                     # the timed region's own figure: its wall time per stage
                     # launch, which also holds the other launches of an
                     # advance and the gaps between launches, so a lower
                     # bound on the kernel's rate (graph replays run the
                     # launches back to back, at a lower shader clock)
                     "timed_region": {
                         "ms_per_launch": wall * 1e3 / (steps * rk),
                         "achieved": algorithmic_bytes_per_step(hydro.nv, rk) * zones * steps / wall / 1e9,
                         "frac": algorithmic_bytes_per_step(hydro.nv, rk) * zones * steps / wall / 1e9
                                 / HBM_PEAK_GBS}},
    }
    g.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(w, u0, args.cpu_seconds, args.cpu_threads or default_cpu_threads())
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
