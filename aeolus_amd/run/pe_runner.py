"""Primitive-equation run driver and per-step diagnostics
(``run/pe/runner.py:55-199``, ``run/pe/diagnostics.py:45-143``).

Fixed user-supplied step (no CFL controller): the integration scheduler is
driven with a constant ceiling, so every step is ``dt_seconds`` except where
count mode clips it to land on an output time or ``t_end``.  Every RK4 stage
and every accepted state is validated (``PrimitiveEquationsModel.validate_state``);
a violation aborts with ``PrimitiveEquationsStateError``.  Written per run
directory: ``pe_coeffs.npy`` (n, 3K+1, L+1, L+1) complex128,
``pe_snapshot_times.npy`` and ``diagnostics/timeseries.csv``.  Figures are not
produced (plotting is outside this package).
"""
from __future__ import annotations

import csv
import math
import pathlib
from typing import Sequence

import numpy as np
import torch

from ..physics.primitive_equations import PrimitiveEquationsModel, PrimitiveEquationsState
from .engine import rk4_step_array
from .schedule import SECONDS_PER_DAY, IntegrationScheduler

__all__ = ["PE_CSV_COLUMNS", "PEDiagnosticsRecorder", "run_pe", "PE_SNAPSHOT_TIMES_FILENAME"]

PE_SNAPSHOT_TIMES_FILENAME = "pe_snapshot_times.npy"
PE_CSV_COLUMNS = ["time_s", "dt_s", "step", "t_min", "t_max", "ps_min", "ps_max", "max_wind_ms",
                  "max_abs_zeta", "max_abs_delta", "max_char_speed_ms", "courant", "total_mass",
                  "mass_rel_drift"]


class PEDiagnosticsRecorder:
    """Append-only CSV of one row per accepted step (``run/pe/diagnostics.py:63-143``).

    Every grid quantity of a row is reduced on the device and the row is one
    host read: temperature extrema over the state and product samplings,
    surface-pressure extrema and the ``p_s`` mass integral by state-grid
    quadrature, max wind speed, max |zeta| and |delta| over all levels.
    The first recorded state fixes the mass baseline.
    """

    def __init__(self, model: PrimitiveEquationsModel, out_dir):
        self.model = model
        self.R = model.R
        self.length_scale = getattr(model.grid, "cfl_length_scale", None)
        w = torch.as_tensor(model.sh.weights, dtype=torch.float64, device=model.device)
        self._area_weights = w * self.R ** 2
        self._mass0: float | None = None
        self.dir = pathlib.Path(out_dir) / "diagnostics"
        self.dir.mkdir(parents=True, exist_ok=True)
        self._csv_file = open(self.dir / "timeseries.csv", "w", newline="", encoding="utf-8")
        self._writer = csv.DictWriter(self._csv_file, fieldnames=PE_CSV_COLUMNS)
        self._writer.writeheader()

    def _device_row(self, state: PrimitiveEquationsState) -> list:
        m = self.model
        K = m.nlev
        y = m._stack(state.coeffs)
        t_mm = m._t_extrema_dev(y)
        ps = torch.exp(m.sh.inv_transform(y[3 * K]))
        u, v = m.wind_on_state_grid(state)
        zd = m.sh.inv_transform(y[:2 * K]).abs()
        vals = torch.stack([t_mm[0], t_mm[1], ps.min(), ps.max(),
                            torch.sqrt(u * u + v * v).max(), zd[:K].max(), zd[K:].max(),
                            torch.sum(self._area_weights * ps)])
        return vals.tolist()

    def record(self, t: float, state: PrimitiveEquationsState, dt: float, step: int) -> dict:
        t_min, t_max, ps_min, ps_max, wind, zmax, dmax, mass = self._device_row(state)
        m = self.model
        max_char = wind + math.sqrt(m.gamma * m.r_dry * max(t_max, 0.0))
        courant = (max_char * dt / self.length_scale
                   if (self.length_scale and dt > 0) else float("nan"))
        if self._mass0 is None:
            self._mass0 = mass
        drift = (mass - self._mass0) / self._mass0 if self._mass0 else 0.0
        row = {"time_s": t, "dt_s": dt, "step": step, "t_min": t_min, "t_max": t_max,
               "ps_min": ps_min, "ps_max": ps_max, "max_wind_ms": wind, "max_abs_zeta": zmax,
               "max_abs_delta": dmax, "max_char_speed_ms": max_char, "courant": courant,
               "total_mass": mass, "mass_rel_drift": drift}
        self._writer.writerow(row)
        self._csv_file.flush()
        return row

    def close(self) -> None:
        if not self._csv_file.closed:
            self._csv_file.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def run_pe(model: PrimitiveEquationsModel, state0: PrimitiveEquationsState, *,
           dt_seconds: float, t_end_days: float, out_dir, snapshot_times: Sequence[float],
           snapshot_mode: str = "count", dt_snapshots: float | None = None) -> int:
    """Integrate the dry primitive equations with a fixed step and persist
    (``run/pe/runner.py:55-199``)."""
    t_end = t_end_days * SECONDS_PER_DAY
    if not (math.isfinite(dt_seconds) and dt_seconds > 0):
        raise ValueError(f"dt_seconds must be finite and > 0, got {dt_seconds}")
    if snapshot_mode not in ("count", "interval"):
        raise ValueError(f"unknown snapshot_mode: {snapshot_mode!r}")
    out_dir = pathlib.Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    model.validate_state(state0, context="initial state")
    state = PrimitiveEquationsState(
        torch.as_tensor(state0.coeffs, device=model.device).to(torch.complex128).clone())
    recorder = PEDiagnosticsRecorder(model, out_dir)
    last_row = recorder.record(0.0, state, dt=0.0, step=0)
    snapshots, stored = [], []
    step = 0

    def validate_stage(ys) -> None:
        model.validate_state(PrimitiveEquationsState(ys),
                             context=f"in an RK4 stage after step {step}")

    sched = IntegrationScheduler(t_end, mode=snapshot_mode, snapshot_times=snapshot_times,
                                 dt_snapshots=dt_snapshots)
    try:
        t = 0.0
        while (ev := sched.next_event(dt_seconds)) is not None:
            kind, dt_step, t_ev = ev
            if kind == "store":
                snapshots.append(state.coeffs.cpu().numpy())
                stored.append(t_ev)
                print(f"Time: {t_ev / 3600.0:8.3f} hrs | Step: {step} ")
                continue
            step += 1
            state = PrimitiveEquationsState(rk4_step_array(model.tendency, state.coeffs, t,
                                                           dt_step,
                                                           stage_validator=validate_stage))
            model.validate_state(state, context=f"at t={t_ev:g} s (step {step})")
            last_row = recorder.record(t_ev, state, dt=dt_step, step=step)
            t = t_ev
    finally:
        recorder.close()
    np.save(out_dir / "pe_coeffs.npy",
            np.stack(snapshots) if snapshots else
            np.empty((0, *state.coeffs.shape), dtype=np.complex128))
    np.save(out_dir / PE_SNAPSHOT_TIMES_FILENAME, np.asarray(stored, dtype=np.float64))
    print(f"Stored {len(stored)} snapshot(s) over {step} fixed {dt_seconds:g} s step(s).")
    print(f"Final diagnostics @ t={last_row['time_s'] / 3600.0:.3f} h: "
          f"T[min,max]=[{last_row['t_min']:.2f}, {last_row['t_max']:.2f}] K, "
          f"p_s[min,max]=[{last_row['ps_min']:.1f}, {last_row['ps_max']:.1f}] Pa, "
          f"max|V|={last_row['max_wind_ms']:.3g} m/s, "
          f"mass drift={last_row['mass_rel_drift']:.2e}")
    return 0
