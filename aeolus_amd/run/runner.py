"""BVE and shallow-water run drivers (``run/bve/runner.py:32-207``,
``run/swe/runner.py:29-135``): state-adaptive advective-CFL integration on
the device with the reference's snapshot contracts and capsule files.

Each accepted step is one fused device RK4 step (BVE) or the RK4 step with
stage validation (shallow water); the step's diagnostics row (one device
read) supplies the max characteristic speed that sets the next ceiling.
Written per run directory:

* BVE: ``vorticity_coeffs.npy`` (n, L+1, L+1) complex128,
  ``vorticity_grid.npy`` (n, N) synthesised vorticity,
  ``bve_snapshot_times.npy``, ``diagnostics/timeseries.csv``,
  ``diagnostics/spectra.npz``;
* SWE: ``swe_coeffs.npy`` (n, 3, L+1, L+1), ``swe_snapshot_times.npy``,
  ``diagnostics/timeseries.csv``.

Figures are not produced (plotting is outside this package).
"""
from __future__ import annotations

import pathlib
from typing import Sequence

import numpy as np
import torch

from ..physics.barotropic import BarotropicState, BarotropicVorticity
from ..physics.shallow_water import ShallowWaterState
from .diagnostics import DiagnosticsRecorder
from .engine import rk4_step, rk4_step_array
from .schedule import (SECONDS_PER_DAY, IntegrationScheduler, advective_cfl_timestep,
                       integrate, validate_snapshot_schedule)
from .swe_diagnostics import SWEDiagnosticsRecorder

__all__ = ["run_bve", "run_swe", "BVE_SNAPSHOT_TIMES_FILENAME", "SWE_SNAPSHOT_TIMES_FILENAME"]

BVE_SNAPSHOT_TIMES_FILENAME = "bve_snapshot_times.npy"
SWE_SNAPSHOT_TIMES_FILENAME = "swe_snapshot_times.npy"


def _mode(snapshot_mode, snapshot_times, dt_snapshots, t_end):
    if snapshot_mode is None:
        snapshot_mode = "interval" if snapshot_times is None else "count"
    if snapshot_mode not in ("count", "interval"):
        raise ValueError(f"unknown snapshot_mode: {snapshot_mode!r}")
    if snapshot_mode == "count":
        if snapshot_times is None:
            raise ValueError("count mode requires an explicit snapshot_times schedule")
        snapshot_times = validate_snapshot_schedule(snapshot_times, t_end)
    elif dt_snapshots is None:
        raise ValueError("interval mode requires dt_snapshots")
    return snapshot_mode, snapshot_times


def run_bve(planet, zeta0_lm, dt_snapshots: float | None, t_end_days: float, out_dir,
            viscosity: float, snapshot_times: Sequence[float] | None = None,
            snapshot_mode: str | None = None) -> int:
    """Integrate the BVE and persist snapshots and diagnostics."""
    out_dir = pathlib.Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    t_end = t_end_days * SECONDS_PER_DAY
    mode, snapshot_times = _mode(snapshot_mode, snapshot_times, dt_snapshots, t_end)
    model = BarotropicVorticity(planet, viscosity=viscosity)
    state = BarotropicState(torch.as_tensor(zeta0_lm, device=planet.sh.device)
                            .to(torch.complex128).clone())
    length_scale = getattr(planet.grid, "cfl_length_scale", None)
    zeta_initial_grid = planet.sh.inv_transform(state.coeffs)
    recorder = DiagnosticsRecorder(planet.sh, planet.so, planet.grid, planet.params.radius,
                                   planet.params.angular_velocity, out_dir)
    first = recorder.record(0.0, state.coeffs, dt=0.0, step=0)
    coeffs, grids, times = [], [], []
    step = 0

    def on_store(t_ev: float) -> None:
        coeffs.append(state.coeffs.cpu().numpy())
        grids.append(planet.sh.inv_transform(state.coeffs).cpu().numpy())
        times.append(t_ev)
        print(f"Time: {t_ev / 3600.0:8.2f} hrs | Step: {step} ")

    def on_step(t0: float, t1: float, dt: float, k: int) -> float:
        nonlocal state, step
        state = rk4_step(model, state, t0, dt)
        step = k
        return recorder.record(t1, state.coeffs, dt=dt, step=k)["max_speed_ms"]

    sched = IntegrationScheduler(t_end, mode=mode, snapshot_times=snapshot_times,
                                 dt_snapshots=dt_snapshots)
    try:
        integrate(sched, advective_cfl_timestep(length_scale, first["max_speed_ms"]),
                  length_scale, on_step=on_step, on_store=on_store)
    finally:
        recorder.close()
    n = state.coeffs.shape[-1]
    np.save(out_dir / "vorticity_coeffs.npy",
            np.stack(coeffs) if coeffs else np.empty((0, n, n), dtype=np.complex128))
    np.save(out_dir / "vorticity_grid.npy",
            np.stack(grids) if grids else np.empty((0, *zeta_initial_grid.shape)))
    np.save(out_dir / BVE_SNAPSHOT_TIMES_FILENAME, np.asarray(times, dtype=np.float64))
    return 0


def run_swe(model, state0: ShallowWaterState, dt_snapshots: float | None, t_end_days: float,
            out_dir, snapshot_times: Sequence[float] | None = None,
            snapshot_mode: str | None = None) -> int:
    """Integrate the shallow-water equations and persist snapshots and
    diagnostics; the initial state, every RK4 stage and every accepted state
    are validated."""
    out_dir = pathlib.Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    t_end = t_end_days * SECONDS_PER_DAY
    mode, snapshot_times = _mode(snapshot_mode, snapshot_times, dt_snapshots, t_end)
    model.validate_state(state0, context="initial state")
    state = ShallowWaterState(torch.as_tensor(state0.coeffs, device=model.device)
                              .to(torch.complex128).clone())
    length_scale = getattr(model.grid, "cfl_length_scale", None)
    recorder = SWEDiagnosticsRecorder(model, out_dir)
    first = recorder.record(0.0, state, dt=0.0, step=0)
    snaps, times = [], []
    step = 0

    def on_store(t_ev: float) -> None:
        snaps.append(state.coeffs.cpu().numpy())
        times.append(t_ev)
        print(f"Time: {t_ev / 3600.0:8.2f} hrs | Step: {step} ")

    def validate_stage(ys) -> None:
        model.validate_state(ShallowWaterState(ys), context=f"in an RK4 stage after step {step}")

    def on_step(t0: float, t1: float, dt: float, k: int) -> float:
        nonlocal state, step
        state = ShallowWaterState(rk4_step_array(model.tendency, state.coeffs, t0, dt,
                                                 stage_validator=validate_stage))
        step = k
        model.validate_state(state, context=f"at t={t1:g} s (step {k})")
        return recorder.record(t1, state, dt=dt, step=k)["max_char_speed_ms"]

    sched = IntegrationScheduler(t_end, mode=mode, snapshot_times=snapshot_times,
                                 dt_snapshots=dt_snapshots)
    try:
        integrate(sched, advective_cfl_timestep(length_scale, first["max_char_speed_ms"]),
                  length_scale, on_step=on_step, on_store=on_store)
    finally:
        recorder.close()
    np.save(out_dir / "swe_coeffs.npy",
            np.stack(snaps) if snaps else np.empty((0, *state.coeffs.shape),
                                                    dtype=np.complex128))
    np.save(out_dir / SWE_SNAPSHOT_TIMES_FILENAME, np.asarray(times, dtype=np.float64))
    return 0
