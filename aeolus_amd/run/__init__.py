"""RK4 steppers, diagnostics and seeded initial conditions."""
from .diagnostics import grid_diagnostics, spectral_diagnostics
from .engine import integrate_fixed, rk4_step, rk4_step_array

__all__ = ["grid_diagnostics", "spectral_diagnostics", "integrate_fixed", "rk4_step",
           "rk4_step_array"]
