"""Barotropic vorticity model on the HIP device (``phy/barotropic.py:8-167``).

``dzeta/dt = -J(psi, zeta + f) + nu lap' zeta + forcing`` with ``psi = zeta/lap``
(l >= 1), ``f`` exact in spectral space (only (1, 0) = 2 Omega sqrt(4 pi/3)),
the Jacobian analysed on the backend's product space and truncated once,
and the l = 0 row pinned to zero.  The whole tendency is one fused library
call (Legendre synthesis, FFT + product + FFT, Legendre analysis with the
tendency epilogue).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import torch

from .. import _native as nat

__all__ = ["BarotropicState", "BarotropicVorticity", "BarotropicBatch"]


@dataclass
class BarotropicState:
    coeffs: torch.Tensor
    tendency: torch.Tensor = None


class BarotropicVorticity:
    def __init__(self, planet, scenario: str = "two_vortices", viscosity: float = 0.0):
        self.planet = planet
        self.nu = float(viscosity)
        self.sh = planet.sh
        self.so = planet.so
        self.grid = planet.grid
        self.R = float(planet.params.radius)
        self.Omega = float(planet.params.angular_velocity)
        dev = self.sh.device
        self.device = dev
        lat = torch.as_tensor(self.grid.point_latitudes, device=dev)
        f = 2 * self.Omega * torch.sin(lat)
        shape = getattr(self.grid, "grid_shape", None)
        if shape is not None and len(shape) == 2 and f.numel() == shape[0] * shape[1]:
            f = f.reshape(shape)
        self.f = f
        L = self.sh.l_max
        self.f_lm = torch.zeros((L + 1, L + 1), dtype=torch.complex128, device=dev)
        self.f_lm[1, 0] = 2.0 * self.Omega * math.sqrt(4.0 * math.pi / 3.0)
        l = torch.arange(L + 1, dtype=torch.float64, device=dev)
        self.laplacian_eig = -l * (l + 1) / self.R ** 2
        self.laplacian_eig[0] = 1.0 / self.R ** 2
        self._ps = self.so.backend.product_space(self.so.product_quadrature)
        self._model = self._create()

    def _create(self):
        handle = ctypes.c_void_p()
        nat.check(nat.lib().aeo_bve_create(self._ps.sh._plan, self.R, self.Omega, self.nu,
                                           ctypes.byref(handle)), "BarotropicVorticity")
        return handle

    def __del__(self):
        h = getattr(self, "_model", None)
        if h is not None and h.value:
            try:
                nat.lib().aeo_model_destroy(h)
            except Exception:
                pass
            self._model = ctypes.c_void_p()

    def _coeffs(self, c) -> torch.Tensor:
        return torch.as_tensor(c, device=self.device).to(torch.complex128).contiguous()

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
        """Members per workgroup of the fused stage kernel (1, 2, 4, 8), or 0 for
        the member-batched GEMM form of the split-precision stage."""
        nat.check(nat.lib().aeo_model_set_tiles(self._model, int(members_per_block)), "set_tiles")

    def set_gemm_tiles(self, syn_cols: int = 64, epi_members: int = 8,
                       syn_ntiles: int = 4) -> None:
        """Tile shape of the GEMM form: synthesis columns per workgroup (64,
        128), members per epilogue workgroup (4, 8) and 16-column MFMA tiles
        per synthesis wave (2, 4)."""
        nat.check(nat.lib().aeo_model_set_gemm_tiles(self._model, int(syn_cols),
                                                     int(epi_members), int(syn_ntiles)),
                  "set_gemm_tiles")

    def vorticity_to_streamfunction(self, vrt_state: BarotropicState) -> torch.Tensor:
        assert isinstance(vrt_state, BarotropicState)
        z = self._coeffs(vrt_state.coeffs)
        out = torch.empty_like(z)
        nat.check(nat.lib().aeo_spec_op(nat.OP_INVLAP_ZERO, self.sh.l_max, self.R, z.data_ptr(),
                                        out.data_ptr(), 1, nat.stream()),
                  "vorticity_to_streamfunction")
        return out

    def streamfunction_to_vorticity(self, psi_coeffs) -> BarotropicState:
        z = torch.zeros_like(self._coeffs(psi_coeffs))
        z[1:, :] = self._coeffs(psi_coeffs)[1:, :] * self.laplacian_eig[1:, None]
        return BarotropicState(z)

    def tendency(self, vrt_state: BarotropicState, forcing_coeffs=None) -> torch.Tensor:
        z = self._coeffs(vrt_state.coeffs)
        out = torch.empty_like(z)
        frc = None if forcing_coeffs is None else self._coeffs(forcing_coeffs)
        nat.check(nat.lib().aeo_model_tendency(self._model, z.data_ptr(),
                                               None if frc is None else frc.data_ptr(),
                                               out.data_ptr(), nat.stream()), "tendency")
        return out

    def rk4(self, coeffs, dt: float, nsteps: int = 1, forcing_coeffs=None,
            out: torch.Tensor | None = None) -> torch.Tensor:
        """``nsteps`` fused RK4 steps (one library call); returns the new coefficients."""
        z = self._coeffs(coeffs)
        y = torch.empty_like(z) if out is None else out
        frc = None if forcing_coeffs is None else self._coeffs(forcing_coeffs)
        nat.check(nat.lib().aeo_model_rk4(self._model, z.data_ptr(), float(dt), int(nsteps),
                                          None if frc is None else frc.data_ptr(),
                                          y.data_ptr(), nat.stream()), "rk4")
        return y

    # -- adjoint (no reference counterpart) ----------------------------------
    def enable_adjoint(self) -> None:
        """Allocate the adjoint workspaces (once); the adjoint calls need them
        and allocate nothing themselves."""
        if not getattr(self, "_adjoint_ready", False):
            nat.check(nat.lib().aeo_model_enable_adjoint(self._model), "enable_adjoint")
            self._adjoint_ready = True

    def checkpoint_elems(self, nsteps: int) -> int:
        """complex128 entries of the stage checkpoints of an ``nsteps`` run."""
        n = ctypes.c_longlong()
        nat.check(nat.lib().aeo_model_checkpoint_elems(self._model, int(nsteps), ctypes.byref(n)),
                  "checkpoint_elems")
        return int(n.value)

    def rk4_record(self, coeffs, dt: float, nsteps: int = 1, forcing_coeffs=None,
                   out: torch.Tensor | None = None):
        """:meth:`rk4` (bit-identical result) that also keeps the input state of
        every RK4 stage; returns ``(y, checkpoints)`` for :meth:`rk4_adjoint`."""
        self.enable_adjoint()
        z = self._coeffs(coeffs)
        y = torch.empty_like(z) if out is None else out
        ck = torch.empty(max(self.checkpoint_elems(nsteps), 1), dtype=torch.complex128,
                         device=self.device)
        frc = None if forcing_coeffs is None else self._coeffs(forcing_coeffs)
        nat.check(nat.lib().aeo_model_rk4_record(self._model, z.data_ptr(), float(dt), int(nsteps),
                                                 None if frc is None else frc.data_ptr(),
                                                 y.data_ptr(), ck.data_ptr(), nat.stream()),
                  "rk4_record")
        return y, ck

    def rk4_adjoint(self, checkpoints: torch.Tensor, dt: float, nsteps: int, ybar_out,
                    out: torch.Tensor | None = None) -> torch.Tensor:
        """Cotangent of the initial state of a recorded run from the cotangent
        ``ybar_out`` of its final state (real inner product of the coefficient
        arrays, the torch.autograd convention for complex tensors)."""
        if checkpoints.numel() < self.checkpoint_elems(nsteps):
            raise ValueError("rk4_adjoint: checkpoint buffer is smaller than the recorded run")
        self.enable_adjoint()
        g = self._coeffs(ybar_out)
        res = torch.empty_like(g) if out is None else out
        nat.check(nat.lib().aeo_model_rk4_adjoint(self._model, checkpoints.data_ptr(), float(dt),
                                                  int(nsteps), g.data_ptr(), res.data_ptr(),
                                                  nat.stream()), "rk4_adjoint")
        return res

    def tendency_adjoint(self, coeffs, kbar) -> torch.Tensor:
        """``T'(coeffs)^T kbar`` for one tendency evaluation."""
        self.enable_adjoint()
        z = self._coeffs(coeffs)
        k = self._coeffs(kbar)
        out = torch.empty_like(z)
        nat.check(nat.lib().aeo_model_tendency_adjoint(self._model, z.data_ptr(), k.data_ptr(),
                                                       out.data_ptr(), nat.stream()),
                  "tendency_adjoint")
        return out


class BarotropicBatch(BarotropicVorticity):
    """``nmembers`` independent BVE states on one Gauss product space, stepped
    together: states are ``(nmembers, L+1, L+1)`` stacks and every kernel of
    the step covers all members (no reference counterpart beyond running
    ``nmembers`` ``BarotropicVorticity`` models in lockstep)."""

    def __init__(self, planet, nmembers: int, viscosity: float = 0.0):
        self.nmembers = int(nmembers)
        super().__init__(planet, viscosity=viscosity)

    def _create(self):
        handle = ctypes.c_void_p()
        nat.check(nat.lib().aeo_bve_create_batch(self._ps.sh._plan, self.R, self.Omega, self.nu,
                                                 self.nmembers, ctypes.byref(handle)),
                  "BarotropicBatch")
        return handle
