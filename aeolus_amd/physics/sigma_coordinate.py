"""Lorenz-staggered sigma-coordinate grid metadata (``phy/sigma_coordinate.py:38-148``).

``SigmaGrid`` holds the K + 1 interfaces ``sigma_{k+1/2}`` (top to bottom,
exactly 0.0 first and 1.0 last, strictly increasing) and the derived
constants the column physics of the primitive-equation tendency uses: full
levels, thicknesses, interface log ratios (``+inf`` for the top layer, never
multiplied into a result) and the Simmons-Burridge ``alpha_k`` (``ln 2`` on
top).  The column operators themselves run on the device inside
``aeo_pe_tendency``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from functools import cached_property

import numpy as np

__all__ = ["SigmaGrid", "SigmaGridError"]


class SigmaGridError(ValueError):
    """The vertical-grid metadata violated a hard structural constraint."""


@dataclass(frozen=True)
class SigmaGrid:
    interfaces: tuple

    def __post_init__(self) -> None:
        try:
            s = tuple(float(x) for x in self.interfaces)
        except (TypeError, ValueError):
            raise SigmaGridError(f"sigma interfaces must be real numbers, got "
                                 f"{self.interfaces!r}")
        if len(s) < 2:
            raise SigmaGridError(f"need at least 2 interfaces (1 layer), got {len(s)}")
        for i, x in enumerate(s):
            if not math.isfinite(x):
                raise SigmaGridError(f"interfaces[{i}] = {x} is not finite")
        if s[0] != 0.0:
            raise SigmaGridError(f"top interface must be exactly 0.0, got {s[0]!r}")
        if s[-1] != 1.0:
            raise SigmaGridError(f"bottom interface must be exactly 1.0, got {s[-1]!r}")
        for i in range(len(s) - 1):
            if not s[i] < s[i + 1]:
                raise SigmaGridError(f"sigma interfaces must be strictly increasing; got "
                                     f"interfaces[{i}] = {s[i]} >= interfaces[{i + 1}] = "
                                     f"{s[i + 1]}")
        object.__setattr__(self, "interfaces", s)

    @classmethod
    def uniform(cls, nlev: int) -> "SigmaGrid":
        """Equal thicknesses: interfaces k / nlev."""
        if not isinstance(nlev, int) or isinstance(nlev, bool) or nlev < 1:
            raise SigmaGridError(f"nlev must be a positive int, got {nlev!r}")
        return cls(tuple(k / nlev for k in range(nlev + 1)))

    @property
    def nlev(self) -> int:
        return len(self.interfaces) - 1

    @cached_property
    def full_levels(self) -> tuple:
        s = self.interfaces
        return tuple(0.5 * (s[k] + s[k + 1]) for k in range(self.nlev))

    @cached_property
    def thickness(self) -> tuple:
        s = self.interfaces
        return tuple(s[k + 1] - s[k] for k in range(self.nlev))

    @cached_property
    def interface_log_ratios(self) -> tuple:
        s = self.interfaces
        return (math.inf,) + tuple(math.log(s[k + 1] / s[k]) for k in range(1, self.nlev))

    @cached_property
    def alpha(self) -> tuple:
        s = self.interfaces
        return (math.log(2.0),) + tuple(
            1.0 - (s[k] / self.thickness[k]) * self.interface_log_ratios[k]
            for k in range(1, self.nlev))

    def interfaces_array(self) -> np.ndarray:
        return np.asarray(self.interfaces, dtype=np.float64)

    def full_levels_array(self) -> np.ndarray:
        return np.asarray(self.full_levels, dtype=np.float64)

    def thickness_array(self) -> np.ndarray:
        return np.asarray(self.thickness, dtype=np.float64)
