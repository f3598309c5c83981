"""Uneven-L BVE ensembles in one graph-captured step.

``Ensemble([(planet_L21, 64), (planet_L42, 64), ...])`` holds one member
batch (``BarotropicBatch``) per L class and steps all of them with
``aeo_ensemble_rk4``: two RK4 steps of every class are captured once into
one hipGraph (one concurrent branch per class) and replayed.  Tile shapes
come from the autotuner.
"""
from __future__ import annotations

import ctypes

import torch

from .. import _native as nat
from ..autotune import tune_members_per_block
from ..physics.barotropic import BarotropicBatch

__all__ = ["Ensemble"]


class Ensemble:
    def __init__(self, groups, viscosity: float = 0.0, precision: int = 0):
        self.models = [BarotropicBatch(pl, int(n), viscosity=viscosity) for pl, n in groups]
        for m in self.models:
            m.set_precision(precision)
        arr = (ctypes.c_void_p * len(self.models))(*[m._model.value for m in self.models])
        handle = ctypes.c_void_p()
        nat.check(nat.lib().aeo_ensemble_create(len(self.models), arr, ctypes.byref(handle)),
                  "Ensemble")
        self._ens = handle
        self.tuning = None

    def __del__(self):
        h = getattr(self, "_ens", None)
        if h is not None and h.value:
            try:
                nat.lib().aeo_ensemble_destroy(h)
            except Exception:
                pass
            self._ens = ctypes.c_void_p()

    def autotune(self, states, dt: float, steps: int = 3, use_cache: bool = True) -> list:
        self.tuning = [tune_members_per_block(m, y, dt, steps=steps, use_cache=use_cache)
                       for m, y in zip(self.models, states)]
        return self.tuning

    def rk4(self, states, dt: float, nsteps: int, out=None, use_graph: bool = True):
        ys = [m._coeffs(y) for m, y in zip(self.models, states)]
        outs = out if out is not None else [torch.empty_like(y) for y in ys]
        yin = (ctypes.c_void_p * len(ys))(*[y.data_ptr() for y in ys])
        yout = (ctypes.c_void_p * len(outs))(*[o.data_ptr() for o in outs])
        nat.check(nat.lib().aeo_ensemble_rk4(self._ens, yin, float(dt), int(nsteps), yout,
                                             1 if use_graph else 0, nat.stream()), "Ensemble.rk4")
        return outs
