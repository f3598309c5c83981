"""RK4 steppers (``run/engine.py:406-431``, ``run/bve/runner.py:211-216``).

When the tendency belongs to one of this package's models, the whole step
is one fused library call (``aeo_model_rk4``).  Any other tendency callable,
or a stage validator, takes the composed path: tendency calls plus device
tensor arithmetic in the reference's order.
"""
from __future__ import annotations

from ..physics.barotropic import BarotropicState, BarotropicVorticity
from ..physics.shallow_water import ShallowWaterModel

__all__ = ["rk4_step", "rk4_step_array", "integrate_fixed"]


def rk4_step(model, y: BarotropicState, t: float, dt: float, forcing_coeffs=None
             ) -> BarotropicState:
    """Classical RK4 for the BVE state (``run/bve/runner.py:211-216``)."""
    if isinstance(model, BarotropicVorticity):
        return BarotropicState(model.rk4(y.coeffs, dt, 1, forcing_coeffs))
    k1 = model.tendency(y, forcing_coeffs)
    k2 = model.tendency(BarotropicState(y.coeffs + 0.5 * dt * k1), forcing_coeffs)
    k3 = model.tendency(BarotropicState(y.coeffs + 0.5 * dt * k2), forcing_coeffs)
    k4 = model.tendency(BarotropicState(y.coeffs + dt * k3), forcing_coeffs)
    return BarotropicState(y.coeffs + (dt / 6.0) * (k1 + 2 * k2 + 2 * k3 + k4))


def rk4_step_array(tendency, y, t: float, dt: float, *, stage_validator=None):
    """Classical RK4 on a stacked coefficient array (``run/engine.py:406-431``)."""
    owner = getattr(tendency, "__self__", None)
    if (stage_validator is None and isinstance(owner, ShallowWaterModel)
            and getattr(tendency, "__func__", None) is ShallowWaterModel.tendency):
        return owner.rk4(y, dt, 1)

    def _stage(ys):
        if stage_validator is not None:
            stage_validator(ys)
        return ys

    k1 = tendency(y)
    k2 = tendency(_stage(y + 0.5 * dt * k1))
    k3 = tendency(_stage(y + 0.5 * dt * k2))
    k4 = tendency(_stage(y + dt * k3))
    return y + (dt / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)


def integrate_fixed(model, coeffs, dt: float, nsteps: int, forcing_coeffs=None):
    """``nsteps`` fixed-dt RK4 steps in one library call (the benchmark loop of
    ``tests/test_backend_parity.py:157-177``)."""
    if isinstance(model, BarotropicVorticity):
        return model.rk4(coeffs, dt, nsteps, forcing_coeffs)
    if isinstance(model, ShallowWaterModel):
        if forcing_coeffs is not None:
            raise ValueError("forcing is only defined for the BVE model")
        return model.rk4(coeffs, dt, nsteps)
    raise TypeError(f"integrate_fixed: unsupported model {type(model).__name__}")
