"""RK4 steppers, integration control, run drivers, diagnostics and seeded
initial conditions."""
from .diagnostics import DiagnosticsRecorder, grid_diagnostics, spectral_diagnostics
from .engine import integrate_fixed, rk4_step, rk4_step_array
from .pe_runner import PE_CSV_COLUMNS, PEDiagnosticsRecorder, run_pe
from .runner import run_bve, run_swe
from .schedule import (IntegrationScheduler, advective_cfl_timestep, count_snapshot_times,
                       integrate, integration_plan, interval_snapshot_times)
from .swe_diagnostics import SWEDiagnosticsRecorder

__all__ = ["grid_diagnostics", "spectral_diagnostics", "integrate_fixed", "rk4_step",
           "rk4_step_array", "DiagnosticsRecorder", "SWEDiagnosticsRecorder", "run_bve",
           "run_swe", "IntegrationScheduler", "advective_cfl_timestep", "count_snapshot_times",
           "interval_snapshot_times", "integrate", "integration_plan", "run_pe",
           "PEDiagnosticsRecorder", "PE_CSV_COLUMNS"]
