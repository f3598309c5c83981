"""Tangent-linear and adjoint restatements of the BVE RK4 step (oracle, test
infrastructure only).

The reference has no adjoint; its only adjoint-flavoured primitive is
``adjoint_sin_theta_d_theta_coeffs`` (``num/spectral_operators.py:243-270``,
restated as :meth:`SpectralOps.sin_theta_d_theta_T`).  These functions
differentiate the oracle's own forward step (:func:`oracle.models.bve_tendency`
after ``phy/barotropic.py:120-167`` and :func:`oracle.models.rk4_step` after
``run/bve/runner.py:211-216``).

Inner product: the real Euclidean product of the dense ``(L+1, L+1)`` complex
arrays, ``<x, y> = sum Re(conj(x) y)``, and the plain sum over grid points for
grid fields.  With ``c_0 = 1`` and ``c_m = 2`` (m > 0) the transposes of the
transforms are
    Syn^T g = c_m * Ana(g / w)          (unweighted analysis),
    Ana^T a = w * Syn(a / c_m)          (synthesis without the factor 2),
for the point weights ``w`` of the transform.

Pinned by tests/test_oracle_adjoint.py: per-operator dot-product identities
(synthesis, analysis, S, d/dlambda, Jacobian), the tangent-linear tendency
against the exact polarisation identity of the quadratic tendency
``T'(y) d = (T(y + d) - T(y - d)) / 2`` and against central differences, the
adjoint tendency and adjoint RK4 step against the tangent-linear ones by
dot-product tests, a multi-step gradient against central finite differences,
and planted mistakes (a dropped c_m factor, S in place of S^T) that the
dot-product tests reject.
"""
from __future__ import annotations

import numpy as np

from .models import Planet, bve_tendency

__all__ = [
    "inner",
    "synthesis_T",
    "analysis_T",
    "d_lambda_T",
    "jacobian_tl",
    "jacobian_T",
    "bve_tendency_tl",
    "bve_tendency_T",
    "rk4_step_record",
    "rk4_step_tl",
    "rk4_step_T",
    "sphere_misfit",
]


def inner(x, y) -> float:
    """``sum Re(conj(x) y)`` over all entries."""
    return float(np.sum(np.real(np.conj(np.asarray(x)) * np.asarray(y))))


def _cm(n: int) -> np.ndarray:
    c = np.full(n, 2.0)
    c[0] = 1.0
    return c


def synthesis_T(sh, g) -> np.ndarray:
    """Transpose of ``sh.inv_transform`` (grid -> dense coefficients)."""
    g = np.asarray(g, dtype=np.float64).ravel()
    a = sh.transform(g / sh.weights)
    return a * _cm(a.shape[1])[None, :]


def analysis_T(sh, a) -> np.ndarray:
    """Transpose of ``sh.transform`` (dense coefficients -> grid)."""
    h = np.array(a, dtype=np.complex128, copy=True)
    h /= _cm(h.shape[1])[None, :]
    return sh.weights * sh.inv_transform(h)


def d_lambda_T(so, a) -> np.ndarray:
    """Transpose of ``d_lambda`` (multiplication by ``i m / R``): ``-i m / R``."""
    return -so.d_lambda(a)


def _fields(so, psh, a):
    lam = psh.inv_transform(so.d_lambda(a))
    s = psh.inv_transform(so.sin_theta_d_theta(a)) / so.R
    return lam, s


def jacobian_tl(so, a, b, da, db, psh, pcoslat):
    """Tangent of the bilinear ``J(a, b)``: ``J(da, b) + J(a, db)``."""
    return so.jacobian(da, b, psh, pcoslat) + so.jacobian(a, db, psh, pcoslat)


def jacobian_T(so, a, b, jbar, psh, pcoslat, dealias: bool = True):
    """``(abar, bbar)``: transpose of ``(da, db) -> J(da, b) + J(a, db)``."""
    a_l, a_s = _fields(so, psh, a)
    b_l, b_s = _fields(so, psh, b)
    jb = so.truncate(jbar) if dealias else np.asarray(jbar, dtype=np.complex128)
    jg = analysis_T(psh, jb) / pcoslat ** 2
    # J = (a_s b_l - a_l b_s) / c^2
    g_as, g_al = jg * b_l, -jg * b_s
    g_bl, g_bs = jg * a_s, -jg * a_l

    def back(g_l, g_s):
        return (d_lambda_T(so, synthesis_T(psh, g_l))
                + so.sin_theta_d_theta_T(synthesis_T(psh, g_s)) / so.R)

    return back(g_al, g_as), back(g_bl, g_bs)


def _lap(planet: Planet):
    l = np.arange(planet.l_max + 1, dtype=np.float64)
    lap = -l * (l + 1.0) / planet.R ** 2
    lap[0] = 1.0 / planet.R ** 2
    return lap


def _psi_eta(planet: Planet, zeta):
    lap = _lap(planet)
    psi = np.zeros_like(zeta)
    psi[1:, :] = zeta[1:, :] / lap[1:, None]
    return psi, zeta + planet.f_lm


def bve_tendency_tl(planet: Planet, zeta, dzeta, nu: float = 0.0) -> np.ndarray:
    """Tangent of :func:`oracle.models.bve_tendency` at ``zeta`` along ``dzeta``."""
    lap = _lap(planet)
    psi, eta = _psi_eta(planet, zeta)
    dpsi = np.zeros_like(dzeta)
    dpsi[1:, :] = dzeta[1:, :] / lap[1:, None]
    out = -jacobian_tl(planet.so, psi, eta, dpsi, dzeta, planet.psh, planet.pcoslat)
    out = out + nu * lap[:, None] * dzeta
    out[0, :] = 0.0
    return out


def bve_tendency_T(planet: Planet, zeta, kbar, nu: float = 0.0) -> np.ndarray:
    """Transpose of the tangent tendency: ``T'(zeta)^T kbar``."""
    lap = _lap(planet)
    kb = np.array(kbar, dtype=np.complex128, copy=True)
    kb[0, :] = 0.0
    psi, eta = _psi_eta(planet, zeta)
    psib, etab = jacobian_T(planet.so, psi, eta, -kb, planet.psh, planet.pcoslat)
    zb = nu * lap[:, None] * kb + etab
    zb[1:, :] += psib[1:, :] / lap[1:, None]
    return zb


def rk4_step_record(tendency, y, dt):
    """``(y_out, [y1, y2, y3, y4])``: one RK4 step (``run/bve/runner.py:211-216``
    arithmetic) and the input state of each stage."""
    k1 = tendency(y)
    y2 = y + 0.5 * dt * k1
    k2 = tendency(y2)
    y3 = y + 0.5 * dt * k2
    k3 = tendency(y3)
    y4 = y + dt * k3
    k4 = tendency(y4)
    return y + (dt / 6.0) * (k1 + 2 * k2 + 2 * k3 + k4), [y, y2, y3, y4]


def rk4_step_tl(tendency, tendency_tl, y, dy, dt):
    """``(y_out, dy_out)``: the step and its tangent along ``dy``."""
    k1 = tendency(y)
    d1 = tendency_tl(y, dy)
    y2, e2 = y + 0.5 * dt * k1, dy + 0.5 * dt * d1
    k2, d2 = tendency(y2), tendency_tl(y2, e2)
    y3, e3 = y + 0.5 * dt * k2, dy + 0.5 * dt * d2
    k3, d3 = tendency(y3), tendency_tl(y3, e3)
    y4, e4 = y + dt * k3, dy + dt * d3
    k4, d4 = tendency(y4), tendency_tl(y4, e4)
    return (y + (dt / 6.0) * (k1 + 2 * k2 + 2 * k3 + k4),
            dy + (dt / 6.0) * (d1 + 2 * d2 + 2 * d3 + d4))


def rk4_step_T(tendency_T, stages, ybar, dt):
    """Transpose of the tangent step at the recorded stage states."""
    y1, y2, y3, y4 = stages
    a = np.asarray(ybar, dtype=np.complex128)
    g4 = tendency_T(y4, (dt / 6.0) * a)
    g3 = tendency_T(y3, (dt / 3.0) * a + dt * g4)
    g2 = tendency_T(y2, (dt / 3.0) * a + 0.5 * dt * g3)
    g1 = tendency_T(y1, (dt / 6.0) * a + 0.5 * dt * g2)
    return a + g4 + g3 + g2 + g1


def sphere_misfit(z, target):
    """``(J, dJ/dz)`` with ``J = 1/2 ||z - target||^2`` in the sphere L2 product
    of real fields: ``sum_l [Re(d_l0)^2 + 2 sum_{0<m<=l} |d_lm|^2] / 2``."""
    d = np.asarray(z) - np.asarray(target)
    n = d.shape[0]
    l = np.arange(n)[:, None]
    m = np.arange(n)[None, :]
    wr = np.where(m <= l, np.where(m == 0, 1.0, 2.0), 0.0)
    wi = np.where((m <= l) & (m > 0), 2.0, 0.0)
    J = 0.5 * float(np.sum(wr * d.real ** 2 + wi * d.imag ** 2))
    return J, wr * d.real + 1j * wi * d.imag


def tendency_funcs(planet: Planet, nu: float = 0.0):
    """``(T, T_tl, T_T)`` closures of the BVE tendency for ``planet``."""
    return (lambda y: bve_tendency(planet, y, nu),
            lambda y, d: bve_tendency_tl(planet, y, d, nu),
            lambda y, k: bve_tendency_T(planet, y, k, nu))
