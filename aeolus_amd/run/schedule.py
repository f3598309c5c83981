"""Model-independent integration control (``run/engine.py:1-403``).

When the solver steps and stores, and nothing about what it integrates:
snapshot schedules (count and legacy interval semantics), the advective CFL
ceiling, the incremental step/store scheduler and the driver loop.  Host
only (stdlib), so every contract is CPU-testable; the device step and the
per-step diagnostics plug in through the ``on_step`` / ``on_store``
callbacks of :func:`integrate`.

The scheduler is an explicit state machine: each :meth:`next_event` call
takes the current ceiling and returns one event, ``("store", 0.0, time)`` or
``("step", dt, t_after)``, with the event sequences of the reference's
coroutine scheduler.
"""
from __future__ import annotations

import math
from typing import Optional, Sequence

__all__ = ["SECONDS_PER_DAY", "CFL_NUMBER", "DEFAULT_CFL_FALLBACK_SECONDS",
           "interval_snapshot_times", "count_snapshot_times", "advective_cfl_timestep",
           "IntegrationScheduler", "integration_plan", "scheduler_tolerance",
           "validate_snapshot_schedule", "integrate"]

SECONDS_PER_DAY = 86400.0
#: advective CFL safety factor, dt = C L / |u| (``run/engine.py:88-91``)
CFL_NUMBER = 0.5
#: ceiling when no CFL condition can be formed (``run/engine.py:93-95``)
DEFAULT_CFL_FALLBACK_SECONDS = 600.0


def _number(name: str, value) -> float:
    try:
        f = float(value)
    except (TypeError, ValueError):
        raise ValueError(f"{name} must be a real number, got {value!r}")
    if not math.isfinite(f):
        raise ValueError(f"{name} must be finite, got {value}")
    return f


def _nonneg(name: str, value) -> float:
    f = _number(name, value)
    if f < 0:
        raise ValueError(f"{name} must be >= 0, got {f}")
    return f


def interval_snapshot_times(dt_snapshots: float, t_end: float) -> list[float]:
    """t = 0 and every interval boundary up to t_end (1e-6 dt tolerance;
    ``run/engine.py:58-72``)."""
    tol = 1e-6 * dt_snapshots
    out: list[float] = []
    k = 0
    while k * dt_snapshots <= t_end + tol:
        out.append(min(k * dt_snapshots, t_end))
        k += 1
    return out


def count_snapshot_times(n_snapshots: int, t_end: float) -> list[float]:
    """N evenly spaced times with exact endpoints: [] / [t_end] / [0 .. t_end]
    (``run/engine.py:75-86``)."""
    if n_snapshots == 0:
        return []
    if n_snapshots == 1:
        return [t_end]
    h = t_end / (n_snapshots - 1)
    out = [i * h for i in range(n_snapshots)]
    out[-1] = t_end
    return out


def advective_cfl_timestep(length_scale: Optional[float], max_speed: float, *,
                           cfl_number: float = CFL_NUMBER,
                           fallback: float = DEFAULT_CFL_FALLBACK_SECONDS) -> float:
    """``cfl_number * length_scale / max_speed``; the fallback for a missing or
    zero length scale or an exactly zero speed; invalid inputs or a
    non-finite / non-positive result raise ``ValueError``
    (``run/engine.py:98-134``)."""
    speed = _nonneg("max_speed", max_speed)
    if length_scale is None:
        return fallback
    scale = _nonneg("length_scale", length_scale)
    if scale == 0.0 or speed == 0.0:
        return fallback
    dt = cfl_number * scale / speed
    if not math.isfinite(dt) or dt <= 0.0:
        raise ValueError(f"advective CFL timestep is not finite and positive (got {dt} from "
                         f"length_scale={scale}, max_speed={speed}, cfl_number={cfl_number})")
    return dt


def _landing_step(t: float, target: float, dt_cfl: float) -> tuple[float, float]:
    """One count-mode step toward `target`: the ceiling, clipped to land on
    the target exactly; a step that cannot advance the clock aborts
    (``run/engine.py:137-163``)."""
    gap = target - t
    if gap <= 0.0:
        raise RuntimeError(f"count planner expected a positive residual, got {gap} at t={t} "
                           f"for target={target}")
    dt = min(dt_cfl, gap)
    if dt == gap:
        return dt, target
    t_after = t + dt
    if t_after <= t:
        raise FloatingPointError(f"count planner step stagnated at t={t}: dt_cfl={dt_cfl} "
                                 f"cannot advance time toward target={target}")
    if t_after > target:
        raise FloatingPointError(f"count planner step overshot target={target}: t={t}, "
                                 f"dt={dt}, t_after={t_after}")
    return dt, t_after


class IntegrationScheduler:
    """Step/store scheduler, one event per call (``run/engine.py:165-290``).

    count mode: every scheduled target (a positive residual is never treated
    as reached) is landed on exactly and stored at its exact time; stepping
    ends at t_end.  interval mode: the legacy countdown -- store at t = 0 and
    at each boundary, stop when the remaining time is within 1e-6 of the
    interval; the final state is stored only for multiples of the interval.
    The ceiling passed with a call sizes the step that call returns; store
    events ignore it.
    """

    def __init__(self, t_end: float, *, mode: str,
                 snapshot_times: Optional[Sequence[float]] = None,
                 dt_snapshots: Optional[float] = None):
        self.mode = mode
        self.t_end = float(t_end)
        self._t = 0.0
        self._done = False
        if mode == "count":
            self._targets = [float(x) for x in (snapshot_times or [])]
            self._i = 0
        elif mode == "interval":
            if dt_snapshots is None:
                raise ValueError("interval mode requires dt_snapshots")
            self._dts = float(dt_snapshots)
            self._tol = 1e-6 * self._dts
            self._to_store = 0.0           # time left to the next boundary
        else:
            raise ValueError(f"unknown integration mode: {mode!r}")

    def next_event(self, dt_cfl: float) -> Optional[tuple]:
        if dt_cfl <= 0 or not math.isfinite(dt_cfl):
            raise ValueError(f"dt_cfl must be finite and positive, got {dt_cfl}")
        if self._done:
            return None
        ev = self._count(dt_cfl) if self.mode == "count" else self._interval(dt_cfl)
        if ev is None:
            self._done = True
        return ev

    def _count(self, dt_cfl: float) -> Optional[tuple]:
        if self._i < len(self._targets) and self._targets[self._i] <= self._t:
            target = self._targets[self._i]
            self._i += 1
            return ("store", 0.0, target)
        if self._i >= len(self._targets) and self._t >= self.t_end:
            return None
        stop = self._targets[self._i] if self._i < len(self._targets) else self.t_end
        dt, self._t = _landing_step(self._t, stop, dt_cfl)
        return ("step", dt, self._t)

    def _interval(self, dt_cfl: float) -> Optional[tuple]:
        if not self._t <= self.t_end + self._tol:
            return None
        if self._to_store <= self._tol:
            self._to_store = self._dts
            return ("store", 0.0, self._t)
        remaining = self.t_end - self._t
        if remaining <= self._tol:
            return None
        dt = min(dt_cfl, self._to_store, remaining)
        if dt <= 0:
            return None
        self._t += dt
        self._to_store = max(0.0, self._to_store - dt)
        return ("step", dt, self._t)


def integration_plan(t_end: float, dt_cfl: float, *, mode: str,
                     snapshot_times: Optional[Sequence[float]] = None,
                     dt_snapshots: Optional[float] = None) -> list[tuple]:
    """Every event for a constant ceiling (``run/engine.py:293-318``)."""
    if dt_cfl <= 0 or not math.isfinite(dt_cfl):
        raise ValueError(f"dt_cfl must be finite and positive, got {dt_cfl}")
    sched = IntegrationScheduler(t_end, mode=mode, snapshot_times=snapshot_times,
                                 dt_snapshots=dt_snapshots)
    events = []
    while (ev := sched.next_event(dt_cfl)) is not None:
        events.append(ev)
    return events


def scheduler_tolerance(t_end: float, times: Sequence[float]) -> float:
    """A fraction of the smallest positive gap and of t_end, at least 1e-12 s
    (``run/engine.py:321-334``)."""
    tol = max(1e-12, 1e-9 * t_end) if t_end > 0 else 1e-12
    gaps = [b - a for a, b in zip(times, list(times)[1:]) if b > a]
    if gaps:
        tol = min(tol, 0.25 * min(gaps))
    return tol


def validate_snapshot_schedule(times: Sequence[float], t_end: float) -> list[float]:
    """Finite, strictly increasing entries in [0, t_end] (1e-9 relative slack
    at the ends, clipped); anything else raises ``ValueError``
    (``run/engine.py:337-362``)."""
    if t_end <= 0 or not math.isfinite(t_end):
        raise ValueError(f"t_end must be finite and positive, got {t_end}")
    slack = max(1e-9, 1e-9 * t_end)
    out: list[float] = []
    for i, t in enumerate(times):
        if not math.isfinite(t):
            raise ValueError(f"snapshot_times[{i}] = {t} is not finite")
        if t < -slack or t > t_end + slack:
            raise ValueError(f"snapshot_times[{i}] = {t} is outside [0, {t_end}]")
        c = max(0.0, min(t_end, float(t)))
        if out and not c > out[-1]:
            raise ValueError(f"snapshot_times must be strictly increasing without duplicates; "
                             f"got {times!r}")
        out.append(c)
    return out


def integrate(scheduler: IntegrationScheduler, dt_cfl: float, length_scale: Optional[float],
              *, on_step, on_store) -> tuple:
    """Drive the scheduler with state-adaptive advective CFL stepping
    (``run/engine.py:365-403``): after each accepted step the ceiling is
    recomputed from the max characteristic speed ``on_step`` returns.
    ``on_step(t_before, t_after, dt, step)``; ``on_store(time)``.  Returns the
    final (t, step)."""
    t, step = 0.0, 0
    while (ev := scheduler.next_event(dt_cfl)) is not None:
        kind, dt, t_ev = ev
        if kind == "store":
            on_store(t_ev)
            continue
        step += 1
        speed = on_step(t, t_ev, dt, step)
        t = t_ev
        dt_cfl = advective_cfl_timestep(length_scale, speed)
    return t, step
