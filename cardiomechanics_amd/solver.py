"""Time loop, adaptive step control and the prescribed-pressure plugin (host side).

  Timing                    CBTiming (CBTiming.cpp:25-162): halve dt on FAILED, double it
                            after Solver.MinSteps successes, stop below Solver.MinTimeStep
  ApplyPressureFromFunction CBApplyPressureFromFunction (CBApplyPressureFromFunction.cpp:20-132):
                            per surface group, piecewise intervals of an analytic function,
                            pressure = Offset + Amplitude * g(t - Offset) on [start, stop + dt/2)
  run                       CBSolver::Run (CBSolver.cpp:1724-1858): a first pass at t = start,
                            then successful steps until StopTime; REPEAT retries, FAILED halves dt;
                            plugins that analyse trial steps (circulation.Circulation) decide
                            whether a converged Newmark trial is accepted
The solver steps themselves (static Newton, Newmark) run on the GPU in libcmg.so.
"""
from __future__ import annotations

import math
import time as _time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .materials import FN_LINEAR, FN_SINUS, analytic


class Timing:
    def __init__(self, start=0.0, stop=1.0, dt=1e-4, min_dt=1e-9, max_dt=None, max_init_dt=None, min_steps=4,
                 fast_relaxation=False):
        self.start, self.stop = start, stop
        self.first_dt = dt
        self.min_dt = min_dt
        self.max_sim_dt = dt if max_dt is None else max_dt
        self.max_init_dt = self.max_sim_dt if max_init_dt is None else max_init_dt
        self.min_steps = min_steps
        self.fast_relaxation = fast_relaxation
        self.reset_for_init()

    def reset_for_init(self):
        self.max_dt = self.max_init_dt
        self.dt = min(self.first_dt, self.max_dt)
        self.counter = 0
        self.lowest_level_counter = 0

    def reset_for_sim(self):
        self.max_dt = self.max_sim_dt
        self.dt = min(self.first_dt, self.max_dt)
        self.counter = 0
        self.lowest_level_counter = 0

    def reset_counter(self):
        self.counter = 0
        self.lowest_level_counter = 0

    def is_valid(self, t):
        return t < self.stop and self.stop - t > 1e-14

    def compute_next(self, status) -> bool:
        """CBTiming::ComputeNextTimeStep (:77-136); False when dt fell below the minimum."""
        if status == _lib.FAILED:
            self.lowest_level_counter = 0
            self.counter *= 2
            self.dt /= 2
            return self.dt >= self.min_dt
        if status == _lib.SUCCESS:
            self.counter += 1
            self.lowest_level_counter += 1
            while (self.counter % 2 == 0 and self.counter % self.min_steps == 0
                   and self.lowest_level_counter >= self.min_steps):
                self.counter //= 2
                if self.dt < self.max_init_dt:
                    self.dt *= 2
                    if not self.fast_relaxation:
                        break
            return True
        return True   # REPEAT: same dt


@dataclass
class Interval:
    kind: int
    start: float
    stop: float
    amplitude: float
    offset: float = 0.0
    invert: bool = False
    fn: list = field(default_factory=lambda: [0.0] * 12)
    prev_time: float = 0.0


class ApplyPressureFromFunction:
    """groups: {group index: [Interval, ...]} (group index = position in the context's group table)."""

    def __init__(self, groups: dict):
        self.groups = groups
        self.pressure = {g: 0.0 for g in groups}

    def apply(self, time, dt):
        for g, intervals in self.groups.items():
            for iv in intervals:
                if iv.start <= time < iv.stop + 0.5 * dt:
                    tt = time - iv.offset                                  # CBDataFromFunction.cpp:45-49
                    f = analytic(iv.kind, iv.fn, tt, iv.prev_time)
                    iv.prev_time = tt
                    self.pressure[g] = iv.offset + iv.amplitude * ((1.0 - f) if iv.invert else f)
                    break
        return self.pressure

    def values(self, n_groups):
        p = np.zeros(n_groups)
        for g, v in self.pressure.items():
            p[g] = v
        return p

    def pressures(self):
        return dict(self.pressure)


def linear_ramp(amplitude, start=0.0, stop=1.0, group=0):
    fn = [start, stop, 1.0, 0, 0, 0, 0, 0, 0, 0, 0, 0]
    return ApplyPressureFromFunction({group: [Interval(FN_LINEAR, start, stop, amplitude, fn=fn)]})


def sinus_pulse(amplitude, start=0.0, stop=0.8, group=0):
    """Amplitude * (1/2 - 1/2 cos(2 pi (t - start)/(stop - start))) on [start, stop]."""
    w = 2 * math.pi / (stop - start)
    fn = [start, stop, 1.0, w, 0.0, 0, 0, 0, 0, 0, 0, 0]
    return ApplyPressureFromFunction({group: [Interval(FN_SINUS, start, stop, amplitude, fn=fn)]})


@dataclass
class StepRecord:
    time: float
    dt: float
    status: int
    newton: int
    krylov: int
    residual: float
    wall: float
    volumes: np.ndarray
    repeats: int = 0          # plugin-requested repeats of this step (circulation coupling iterations - 1)


def run(ctx, kind: str, timing: Timing, plugin=None, on_step=None, warm_start=None, max_steps=None,
        plugins=None, repeat_from_trial=True):
    """CBSolver::Run (CBSolver.cpp:1724-1858) with the GPU solver step.  kind = 'static'
    or 'newmark'.  plugin / plugins: objects with apply(time, dt) and pressures()
    ({pressure group: Pa}); a plugin with analyze(volumes) (the circulation) sees the
    cavity volumes of each converged Newmark trial before it is accepted
    (CBSolverNewmarkBeta.cpp:928-976) and answers through its `status`: REPEAT solves
    the step again with the plugin's next pressure estimate (from the last trial's
    increment when repeat_from_trial, else from the predictor as the reference),
    FAILED rejects it; on a rejected step every plugin not asking for a repeat steps
    back.  warm_start(ctx, t, dt) may return an increment used as the Newton initial
    guess.  Stops early when every plugin's exit_check says so."""
    plugins = list(plugins or []) + ([plugin] if plugin is not None else [])
    analyzers = [p for p in plugins if hasattr(p, "analyze")]
    if analyzers and kind != "newmark":
        raise ValueError("plugins that analyse trial steps need the Newmark solver")
    records = []
    time = timing.start
    first = True
    timing.reset_for_init()
    steps = 0
    if analyzers:
        vols0 = ctx.cavity_volumes()
        for p in analyzers:
            if hasattr(p, "init_volumes"):
                p.init_volumes(vols0, timing.dt)
    while timing.is_valid(time) or first:
        repeat = False
        repeats = 0
        while True:
            t = time + timing.dt
            if first:
                time = timing.start - timing.dt
                t = timing.start
            if plugins:
                pr = np.zeros(ctx.n_groups)
                for p in plugins:
                    p.apply(t, timing.dt)
                    for g, v in p.pressures().items():
                        pr[g] = v
                    if getattr(p, "wants_zero_velocity", False):
                        ctx.set_state(velocity=np.zeros(ctx.n_dofs), acceleration=np.zeros(ctx.n_dofs))
                        p.wants_zero_velocity = False
                ctx.set_pressures(pr)
            if warm_start is not None and not repeat:
                g = warm_start(ctx, t, timing.dt)
                if g is not None:
                    ctx.set_initial_guess(g)
            t0 = _time.perf_counter()
            if kind == "static":
                rc, st = ctx.static_step(t, timing.dt)
            elif not analyzers:
                rc, st = ctx.newmark_step(t, timing.dt)
            else:
                rc, st = ctx.newmark_solve(t, timing.dt, from_trial=repeat and repeat_from_trial)
                if rc == _lib.SUCCESS:
                    vols = ctx.trial_volumes()
                    for p in analyzers:
                        p.analyze(vols)
                    fb = [p.status for p in analyzers]
                    if _lib.FAILED in fb:
                        rc = _lib.FAILED
                    elif _lib.REPEAT in fb:
                        rc = _lib.REPEAT
                    else:
                        ctx.commit_step()
            wall = _time.perf_counter() - t0
            if rc == _lib.FAILED:
                for p in plugins:
                    if hasattr(p, "step_back"):
                        p.step_back()
            elif rc == _lib.REPEAT:
                repeats += 1
                for p in plugins:
                    if hasattr(p, "step_back") and getattr(p, "status", None) != _lib.REPEAT:
                        p.step_back()
            repeat = rc == _lib.REPEAT
            ok = timing.compute_next(rc)
            if rc == _lib.SUCCESS:
                vols = ctx.cavity_volumes()
                rec = StepRecord(t, timing.dt if first else t - time, rc, st.newton_iterations,
                                 st.krylov_iterations, st.residual_norm, wall, vols, repeats)
                records.append(rec)
                if on_step:
                    on_step(rec, st)
                time = t
                break
            if not ok:
                raise RuntimeError(f"minimum time step reached at t={t}: simulation failed (last attempt: "
                                   f"dt={timing.dt:.3e}, Newton reason {st.converged_reason} after "
                                   f"{st.newton_iterations} iterations, ||r0||={st.initial_residual_norm:.3e}, "
                                   f"||r||={st.residual_norm:.3e}, corrupt element {st.corrupt_element})")
        if first:
            first = False
            timing.reset_for_sim()
            timing.reset_counter()
        for p in plugins:
            if hasattr(p, "write"):
                p.write(time)
        exits = [p.exit_check(time) for p in plugins if hasattr(p, "exit_check")]
        steps += 1
        if exits and any(exits):
            break
        if max_steps is not None and steps >= max_steps:
            break
    return records
