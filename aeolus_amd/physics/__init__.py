"""Barotropic-vorticity, shallow-water and primitive-equation models, sigma grids and bottom topography."""
from .barotropic import BarotropicBatch, BarotropicState, BarotropicVorticity
from .primitive_equations import (PrimitiveEquationsModel, PrimitiveEquationsState,
                                  PrimitiveEquationsStateError, isothermal_rest_state)
from .shallow_water import (ShallowWaterModel, ShallowWaterState, ShallowWaterStateError,
                            WILLIAMSON_GRAVITY)
from .sigma_coordinate import SigmaGrid, SigmaGridError
from .topography import Topography, TopographyError

__all__ = ["BarotropicBatch", "BarotropicState", "BarotropicVorticity", "ShallowWaterModel", "ShallowWaterState",
           "ShallowWaterStateError", "WILLIAMSON_GRAVITY", "Topography", "TopographyError",
           "PrimitiveEquationsModel", "PrimitiveEquationsState", "PrimitiveEquationsStateError",
           "isothermal_rest_state", "SigmaGrid", "SigmaGridError"]
