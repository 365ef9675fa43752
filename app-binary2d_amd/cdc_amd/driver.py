"""The run loop around the GPU advance: recurring side effects (iteration
message, time series, progress report, checkpoints) and the fallback stack
with safe mode, as src/main.rs:24-88 (side_effects) and src/main.rs:99-210
(run) do it.

Differences by design: fallback states are device snapshots of the solver
(cdc_snapshot_*), copied at device bandwidth instead of cloned on the host;
the physics for each advance is pushed to the solver only when it changes.
"""
import copy
import math
import os
import time
from dataclasses import dataclass
from typing import Optional

from . import checkpoint as ckpt
from .solver import HydroError, Snapshot

ORBITAL_PERIOD = 2.0 * math.pi


@dataclass
class Control:
    """Control, src/app.rs:94-131 (num_threads has no meaning here)."""
    num_orbits: float
    checkpoint_interval: float = 1.0
    time_series_interval: Optional[float] = None
    fold: int = 10
    num_threads: Optional[int] = None
    output_directory: str = "."
    fallback_stack_size: int = 0


class RecurringTask:
    """RecurringTask, src/tasks.rs:10-30 (count and next_time persist)."""

    def __init__(self, count=0, next_time=0.0):
        self.count = count
        self.next_time = next_time
        self.count_this_run = 0
        self.last_performed = time.perf_counter()

    def advance(self, interval):
        seconds = time.perf_counter() - self.last_performed
        self.count += 1
        self.count_this_run += 1
        self.next_time += interval
        self.last_performed = time.perf_counter()
        return seconds

    def elapsed_hours(self):
        return (time.perf_counter() - self.last_performed) / 3600.0

    def to_dict(self):
        return {"count": self.count, "next_time": self.next_time}


def new_tasks(saved=None):
    saved = saved or {}
    return {name: RecurringTask(**saved.get(name, {})) for name in ckpt.TASKS}


def _clone_tasks(tasks):
    out = {}
    for k, t in tasks.items():
        c = RecurringTask(t.count, t.next_time)
        c.count_this_run, c.last_performed = t.count_this_run, t.last_performed
        out[k] = c
    return out


class Runner:
    """One simulation: a Solver plus its configuration, tasks and time series."""

    def __init__(self, solver, physics, control, model=None, tasks=None, time_series=None, log=print):
        self.solver = solver
        self.physics = physics
        self.control = control
        self.model = model
        self.tasks = tasks if tasks is not None else new_tasks()
        self.time_series = list(time_series or [])
        self.log = log
        self.checkpoints = []
        self._pushed = None

    # ---- side effects (src/main.rs:24-88) --------------------------------
    def side_effects(self, dt):
        s, t, c = self.solver, self.tasks, self.control
        now = s.time()
        if t["iteration_message"].next_time <= now:
            secs = t["iteration_message"].advance(0.0)
            if t["iteration_message"].count_this_run > 1 and dt > 0:
                mzps = 1e-6 * s.mesh.total_zones() / secs * c.fold
                self.log(f"orbit={now / ORBITAL_PERIOD:.5f} steps/orbit={ORBITAL_PERIOD / dt:.2e} Mzps={mzps:.2f}")
        if t["record_time_series"].next_time <= now and c.time_series_interval is not None:
            t["record_time_series"].advance(c.time_series_interval * ORBITAL_PERIOD)
            self.time_series.append(s.time_series_sample())
        if t["report_progress"].next_time <= now:
            if t["report_progress"].count > 0:
                self.log(f"orbits / hour {1.0 / max(t['report_progress'].elapsed_hours(), 1e-12):.2f}")
            t["report_progress"].advance(ORBITAL_PERIOD)
        if t["write_checkpoint"].next_time <= now:
            os.makedirs(c.output_directory, exist_ok=True)
            name = os.path.join(c.output_directory, f"chkpt.{t['write_checkpoint'].count:04d}.cbor")
            if t["write_checkpoint"].count_this_run > 0 or t["write_checkpoint"].count == 0:
                self.write_checkpoint(name)
            t["write_checkpoint"].advance(c.checkpoint_interval * ORBITAL_PERIOD)

    def write_checkpoint(self, path):
        st = self.solver.get_state()
        ckpt.write(path, st, self.solver.mesh, self.solver.hydro, self.physics, model=self.model,
                   control=vars(self.control), tasks={k: v.to_dict() for k, v in self.tasks.items()},
                   time_series=self.time_series, torque=self.solver.torque())
        self.checkpoints.append(path)

    # ---- the loop (src/main.rs:99-210) -----------------------------------
    def _use_physics(self, p):
        if self._pushed is not p:
            self.solver.set_physics(p)
            self._pushed = p

    def run(self):
        c, ph = self.control, self.physics
        safe = ph.safe()
        fallback = []            # [(snapshot, time series, tasks)], oldest first
        spare = []
        crash = None             # time of the last crash
        dt = 0.0
        self.side_effects(dt)
        while self.solver.time() < c.num_orbits * ORBITAL_PERIOD:
            if c.fallback_stack_size > 1:
                if len(fallback) > c.fallback_stack_size:
                    spare.append(fallback.pop(0)[0])
                snap = spare.pop() if spare else Snapshot(self.solver)
                self.solver.save(snap)
                fallback.append((snap, list(self.time_series), _clone_tasks(self.tasks)))
            self._use_physics(safe if crash is not None else ph)
            try:
                dt = self.solver.advance(c.fold)
                if crash is not None and crash + ph.safe_mode_duration * ORBITAL_PERIOD < self.solver.time():
                    crash = None
                    self.log(f"surpassed previous crash time plus a margin of {ph.safe_mode_duration} orbits, "
                             f"proceeding in normal mode")
            except HydroError as e:
                if not fallback or crash is not None:
                    os.makedirs(c.output_directory, exist_ok=True)
                    name = os.path.join(c.output_directory, "chkpt.fail.cbor")
                    if self.tasks["write_checkpoint"].count_this_run > 0 or self.tasks["write_checkpoint"].count == 0:
                        self.write_checkpoint(name)
                    raise
                snap, series, tasks = fallback[0]
                crash = fallback[-1][0].time()
                self.solver.restore(snap)
                spare.extend(f[0] for f in fallback)
                fallback.clear()
                self.log(f"{e}")
                self.log(f"rewind to orbit {self.solver.time() / ORBITAL_PERIOD:.5f}, proceed in safety mode...")
                self.time_series, self.tasks = series, tasks
            self.side_effects(dt)
        for s in spare + [f[0] for f in fallback]:
            s.close()
        return self
