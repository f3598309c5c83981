"""Barotropic-vorticity and shallow-water models and bottom topography."""
from .barotropic import BarotropicBatch, BarotropicState, BarotropicVorticity
from .shallow_water import (ShallowWaterModel, ShallowWaterState, ShallowWaterStateError,
                            WILLIAMSON_GRAVITY)
from .topography import Topography, TopographyError

__all__ = ["BarotropicBatch", "BarotropicState", "BarotropicVorticity", "ShallowWaterModel", "ShallowWaterState",
           "ShallowWaterStateError", "WILLIAMSON_GRAVITY", "Topography", "TopographyError"]
