"""Rotating shallow-water model on the HIP device (``phy/shallow_water.py:95-539``).

State: one ``(3, L+1, L+1)`` complex128 stack (zeta, delta, phi').  The
tendency is one fused library call: 10 Legendre syntheses, 12 inverse FFTs,
the flux products of ``phy/shallow_water.py:377-386`` on the product grid, 4
forward FFTs, 4 truncated Legendre analyses and the assembly epilogue
(``:389-418``) including topography and hyperdiffusion terms.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import torch

from .. import _native as nat
from .topography import Topography

__all__ = ["ShallowWaterModel", "ShallowWaterState", "ShallowWaterStateError",
           "WILLIAMSON_GRAVITY", "ZETA", "DELTA", "PHI", "PROGNOSTICS"]

WILLIAMSON_GRAVITY = 9.80616
ZETA, DELTA, PHI = 0, 1, 2
PROGNOSTICS = ("zeta", "delta", "phi")


class ShallowWaterStateError(ValueError):
    """A shallow-water state violated a hard physical/numerical constraint."""


@dataclass
class ShallowWaterState:
    coeffs: torch.Tensor

    @classmethod
    def from_fields(cls, zeta_lm, delta_lm, phi_lm) -> "ShallowWaterState":
        return cls(torch.stack([torch.as_tensor(x).to(torch.complex128)
                                for x in (zeta_lm, delta_lm, phi_lm)]))

    @classmethod
    def zeros(cls, l_max: int, device=None) -> "ShallowWaterState":
        n = l_max + 1
        return cls(torch.zeros((3, n, n), dtype=torch.complex128, device=device))

    @property
    def zeta(self):
        return self.coeffs[ZETA]

    @property
    def delta(self):
        return self.coeffs[DELTA]

    @property
    def phi(self):
        return self.coeffs[PHI]


class ShallowWaterModel:
    _MONOPOLE_RTOL = 1e-10

    def __init__(self, planet, *, gravity: float = WILLIAMSON_GRAVITY, mean_depth: float,
                 hyperdiffusion_nu4: float = 0.0, topography: Topography | None = None):
        if not (math.isfinite(gravity) and gravity > 0):
            raise ValueError(f"gravity must be finite and > 0, got {gravity}")
        if not (math.isfinite(mean_depth) and mean_depth > 0):
            raise ValueError(f"mean_depth must be finite and > 0, got {mean_depth}")
        if not (math.isfinite(hyperdiffusion_nu4) and hyperdiffusion_nu4 >= 0):
            raise ValueError(f"hyperdiffusion_nu4 must be finite and >= 0, got {hyperdiffusion_nu4}")
        self.planet = planet
        self.sh = planet.sh
        self.so = planet.so
        self.grid = planet.grid
        self.device = self.sh.device
        self.R = float(planet.params.radius)
        self.Omega = float(planet.params.angular_velocity)
        self.gravity = float(gravity)
        self.mean_depth = float(mean_depth)
        self.phi0 = self.gravity * self.mean_depth
        self.nu4 = float(hyperdiffusion_nu4)
        self.l_max = self.sh.l_max
        n = self.l_max + 1
        self.f_lm = torch.zeros((n, n), dtype=torch.complex128, device=self.device)
        self.f_lm[1, 0] = 2.0 * self.Omega * math.sqrt(4.0 * math.pi / 3.0)
        l = torch.arange(n, dtype=torch.float64, device=self.device)
        self.lap_eigs = -l * (l + 1.0) / self.R ** 2
        self._ps = self.so.backend.product_space(self.so.product_quadrature)
        coslat = getattr(self.grid, "coslat", None)
        if coslat is None:
            coslat = torch.cos(torch.as_tensor(self.grid.point_latitudes))
        self._state_coslat = torch.clamp(torch.as_tensor(coslat, device=self.device), min=1e-8)
        self._trunc_cut = (2 * self.l_max) // 3
        if topography is None:
            topography = Topography.flat(self.l_max, device=self.device)
        if topography.l_max != self.l_max:
            raise ValueError(f"topography truncation l_max={topography.l_max} does not match "
                             f"the model truncation l_max={self.l_max}")
        self.topography = topography
        self.has_topography = not topography.is_flat
        if self.has_topography:
            self.phi_s_lm = topography.surface_geopotential_lm(self.gravity).to(self.device)
            self.phi_s_anom_lm = self.phi_s_lm.clone()
            self.phi_s_anom_lm[0, 0] = 0.0
            samplings = [self.sh] + ([self._ps.sh] if self._ps.sh is not self.sh else [])
            lo, hi = math.inf, -math.inf
            for sh in samplings:
                e = topography.elevation_on(sh)
                lo, hi = min(lo, float(e.min())), max(hi, float(e.max()))
            self.surface_elevation_extrema = (lo, hi)
            self._phi_s_state = self.sh.inv_transform(self.phi_s_lm)
            phi_s_ptr = self.phi_s_lm.contiguous()
        else:
            self.phi_s_lm = torch.zeros((n, n), dtype=torch.complex128, device=self.device)
            self.phi_s_anom_lm = self.phi_s_lm.clone()
            self.surface_elevation_extrema = (0.0, 0.0)
            self._phi_s_state = None
            phi_s_ptr = None
        handle = ctypes.c_void_p()
        nat.check(nat.lib().aeo_swe_create(self._ps.sh._plan, self.R, self.Omega, self.gravity,
                                           self.mean_depth, self.nu4,
                                           None if phi_s_ptr is None else phi_s_ptr.data_ptr(),
                                           ctypes.byref(handle)), "ShallowWaterModel")
        self._model = handle

    def __del__(self):
        h = getattr(self, "_model", None)
        if h is not None and h.value:
            try:
                nat.lib().aeo_model_destroy(h)
            except Exception:
                pass
            self._model = ctypes.c_void_p()

    def set_precision(self, nslices: int = 0) -> None:
        """0: FP64 Legendre / Y-basis contractions; k in {4..7}: split INT8 MFMA digits.

        Builds the product plan's digit tables on first use.  Everything
        else of the step stays FP64."""
        k = int(nslices)
        if k:
            sh = self._ps.sh
            if getattr(sh, "nslices", 0) != k:
                sh.enable_split(k)
        nat.check(nat.lib().aeo_model_set_precision(self._model, k), "set_precision")
        self.nslices = k

    def set_tiles(self, members_per_block: int) -> None:
        nat.check(nat.lib().aeo_model_set_tiles(self._model, int(members_per_block)), "set_tiles")

    def _stack(self, coeffs) -> torch.Tensor:
        y = torch.as_tensor(coeffs, device=self.device).to(torch.complex128).contiguous()
        n = self.l_max + 1
        if tuple(y.shape) != (3, n, n):
            raise ValueError(f"shallow-water state must have shape (3, {n}, {n})")
        return y

    # -- Helmholtz / wind (phy/shallow_water.py:286-330) -------------------
    def _inv_laplacian(self, coeffs):
        c = torch.as_tensor(coeffs, device=self.device).to(torch.complex128).contiguous()
        out = torch.empty_like(c)
        nat.check(nat.lib().aeo_spec_op(nat.OP_INVLAP_ZERO, self.l_max, self.R, c.data_ptr(),
                                        out.data_ptr(), c.numel() // (self.l_max + 1) ** 2,
                                        nat.stream()), "inv_laplacian")
        return out

    def helmholtz(self, state: ShallowWaterState):
        return self._inv_laplacian(state.coeffs[ZETA]), self._inv_laplacian(state.coeffs[DELTA])

    def _deriv_fields(self, sh, coeffs):
        lam = sh.inv_transform(self.so.d_lambda_coeffs(coeffs))
        snt = sh.inv_transform(self.so.sin_theta_d_theta_coeffs(coeffs)) / self.R
        return lam, snt

    def _wind_from(self, sh, coslat, psi_lm, chi_lm):
        psi_lam, psi_snt = self._deriv_fields(sh, psi_lm)
        chi_lam, chi_snt = self._deriv_fields(sh, chi_lm)
        return (psi_snt + chi_lam) / coslat, (psi_lam - chi_snt) / coslat

    def wind_on_state_grid(self, state: ShallowWaterState):
        psi, chi = self.helmholtz(state)
        return self._wind_from(self.sh, self._state_coslat, psi, chi)

    def surface_geopotential_on_state_grid(self):
        return self._phi_s_state

    # -- tendencies --------------------------------------------------------
    def tendency(self, coeffs) -> torch.Tensor:
        y = self._stack(coeffs)
        out = torch.empty_like(y)
        nat.check(nat.lib().aeo_model_tendency(self._model, y.data_ptr(), None, out.data_ptr(),
                                               nat.stream()), "tendency")
        return out

    def tendency_state(self, state: ShallowWaterState) -> ShallowWaterState:
        return ShallowWaterState(self.tendency(state.coeffs))

    def rk4(self, coeffs, dt: float, nsteps: int = 1, out: torch.Tensor | None = None):
        """``nsteps`` fused RK4 steps without stage validation (one library call)."""
        y = self._stack(coeffs)
        res = torch.empty_like(y) if out is None else out
        nat.check(nat.lib().aeo_model_rk4(self._model, y.data_ptr(), float(dt), int(nsteps), None,
                                          res.data_ptr(), nat.stream()), "rk4")
        return res

    # -- validation (phy/shallow_water.py:428-539) -------------------------
    def _stats(self, state: ShallowWaterState, w=None):
        """Device statistics of one step (one host read): the SWE grid stats of
        the state sampling (aeo_grid_stats) and the min/max of phi on the
        product sampling when it differs."""
        from ..run.diagnostics import grid_stats
        phi_c = state.coeffs[PHI]
        u, v = self.wind_on_state_grid(state)
        phi_g = self.sh.inv_transform(phi_c)
        mean_phi_s = (float(self.phi_s_lm[0, 0].real) / math.sqrt(4.0 * math.pi)
                      if self._phi_s_state is not None else 0.0)
        w = torch.as_tensor(self.sh.weights, dtype=torch.float64) if w is None else w
        gs = grid_stats(nat.GRID_STATS_SWE, w, u, v, phi_g, self._phi_s_state,
                        k=(self.phi0, mean_phi_s, self.gravity))
        if self._ps.sh is not self.sh:
            mm = grid_stats(nat.GRID_STATS_MINMAX, None, self._ps.sh.inv_transform(phi_c))[:2]
        else:
            mm = gs[1:3] - self.phi0
        return u, v, phi_g, gs, mm

    def characteristic_fields(self, state: ShallowWaterState) -> dict:
        """State-grid u, v, wind speed and total geopotential, and the
        total-geopotential extrema over every sampling
        (``phy/shallow_water.py:456-468``)."""
        u, v, phi_g, gs, mm = self._stats(state)
        h = torch.cat([gs[1:3], mm]).cpu().tolist()
        ext = (min(h[0], self.phi0 + h[2]), max(h[1], self.phi0 + h[3]))
        return {"u": u, "v": v, "wind_speed": torch.sqrt(u * u + v * v),
                "phi_total": self.phi0 + phi_g, "phi_total_extrema": ext}

    def max_characteristic_speed(self, state: ShallowWaterState) -> float:
        """max|u| + sqrt(max(Phi0 + phi)) over every sampling, clamped at 0
        inside the root (``phy/shallow_water.py:470-484``); one host read."""
        _, _, _, gs, mm = self._stats(state)
        h = torch.cat([gs[0:3], mm]).cpu().tolist()
        phi_max = max(h[2], self.phi0 + h[4])
        return h[0] + math.sqrt(max(phi_max, 0.0))

    def _phi_extrema_dev(self, phi_c) -> torch.Tensor:
        """Device ``[min, max]`` of phi over the state and product samplings."""
        from ..run.diagnostics import grid_stats
        samplings = [self.sh] + ([self._ps.sh] if self._ps.sh is not self.sh else [])
        mm = [grid_stats(nat.GRID_STATS_MINMAX, None, sh.inv_transform(phi_c))[:2]
              for sh in samplings]
        return torch.stack([torch.stack([x[0] for x in mm]).min(),
                            torch.stack([x[1] for x in mm]).max()])

    def total_geopotential_extrema(self, state: ShallowWaterState) -> tuple[float, float]:
        """(min, max) of Phi0 + phi over every sampling (``:494-539``); one host read."""
        lo, hi = self._phi_extrema_dev(self._stack(state.coeffs)[PHI]).tolist()
        return self.phi0 + lo, self.phi0 + hi

    def validate_state(self, state: ShallowWaterState, *, context: str = "") -> None:
        """Finiteness, per-prognostic monopoles and strictly positive thickness
        over every sampling (``:428-454``).  Every check is reduced on the
        device; the outcome is one host read."""
        where = f" {context}" if context else ""
        coeffs = self._stack(state.coeffs)
        finite = torch.isfinite(coeffs).all()
        mono = coeffs[:, 0, 0].abs()
        scale = torch.linalg.norm(coeffs.reshape(3, -1), dim=1)
        lo_dev = self._phi_extrema_dev(coeffs[PHI])[0]
        host = torch.cat([finite[None].to(torch.float64), mono, scale, lo_dev[None]]).tolist()
        if host[0] != 1.0:
            raise ShallowWaterStateError(f"shallow-water state contains NaN/Inf coefficients{where}")
        for idx, name in enumerate(PROGNOSTICS):
            m, sc = host[1 + idx], host[4 + idx]
            if m > self._MONOPOLE_RTOL * max(sc, 1e-30):
                raise ShallowWaterStateError(
                    f"{name} monopole is nonzero{where}: |a00| = {m:g} (field norm {sc:g})")
        lo = self.phi0 + host[7]
        if not lo > 0.0:
            raise ShallowWaterStateError(
                f"fluid thickness is not strictly positive{where}: min(Phi0 + phi) = {lo:g} "
                f"m^2/s^2 over the state/product samplings (Phi0 = {self.phi0:g})")
