"""Primitive-equation restatements (oracle, test infrastructure only).

* :class:`Sigma` restates the Lorenz-staggered grid metadata of
  ``phy/sigma_coordinate.py:43-148`` (full levels, thicknesses, interface log
  ratios with +inf on top, Simmons-Burridge ``alpha`` with ``ln 2`` on top).
* Column operators: :func:`hydrostatic` (``phy/sigma_coordinate.py:164-209``),
  :func:`column_mass_tendency` / :func:`interface_sigma_dot`
  (``:212-252``), :func:`vertical_advection` (``:280-305``),
  :func:`omega_over_p` (``:373-398``), :func:`layer_mass_residual`
  (``:462-479``), :func:`energy_exchange` (``:401-459``) and
  :func:`vertical_sbp` (``:338-370``).
* :func:`vector_curl_div` restates the weak-form vector analysis of
  ``num/spectral_operators.py:399-502``: half-metric analyses at degree L+1
  on the product sampling, ``div = (i m/R) a + (S^T_ext b)/R``,
  ``curl = (i m/R) b - (S^T_ext a)/R`` with the unclipped ``C+`` at l = L.
* :func:`pe_tendency` restates ``PrimitiveEquationsModel.tendency``
  (``phy/primitive_equations.py:482-753``): product-grid reconstruction,
  thermodynamic and ln(p_s) tendencies, vector-invariant momentum tendencies
  with the spectral hydrostatic geopotential; :func:`isothermal_rest_state`
  follows ``:139-161``; :func:`state_wind`, :func:`temperature_extrema` and
  :func:`max_characteristic_speed` follow ``:280-307, 764-806``.

Pinned by (tests/test_oracle_pe.py): the reference's sigma-coordinate hand
values (``tests/test_sigma_coordinate.py:76-242, 355-441``: alpha and log
ratios, isothermal interface/full-level geopotential closed forms, single
layer sigma_dot, uniform-G omega/p, K = 2 advection, linear-profile closed
form), round-off closures (layer mass, SBP, energy exchange), the weak-form
closed forms ``curl F = lap psi``, ``div F = lap chi`` on the Gauss 3/2-rule
grid to 1e-12 (``tests/test_pe_tendency.py:439-493``), and the tendency's
exact properties: isothermal rest is bitwise zero, BVE degeneracy against
:func:`oracle.models.bve_tendency` to 1e-12, zero flow with structured T
gives ``-lap Phi`` (``tests/test_pe_tendency.py:802-1000``); the resting
state's characteristic speed is the Lamb bound ``sqrt(gamma R_d T0)``
(``tests/test_primitive_equations.py:428-441``) and a solid-body rotation
``psi = -a R^2 Omega' sin(lat)`` has ``u = a Omega' R cos(lat)``, ``v = 0``.
"""
from __future__ import annotations

import math

import numpy as np

from .grids import GaussGrid
from .transforms import DenseTransform, GaussTransform

__all__ = ["R_DRY", "CP_DRY", "Sigma", "hydrostatic", "column_mass_tendency",
           "interface_sigma_dot", "vertical_advection", "omega_over_p",
           "layer_mass_residual", "energy_exchange", "vertical_sbp", "extended_transform",
           "vector_curl_div", "isothermal_rest_state", "PEModel", "pe_tendency", "state_wind",
           "temperature_extrema", "max_characteristic_speed"]

R_DRY = 287.04          # phy/primitive_equations.py:52
CP_DRY = 1004.64        # phy/primitive_equations.py:53


class Sigma:
    """Sigma-grid metadata (``phy/sigma_coordinate.py:43-148``)."""

    def __init__(self, interfaces):
        s = tuple(float(x) for x in interfaces)
        if len(s) < 2 or s[0] != 0.0 or s[-1] != 1.0 or \
                any(not s[i] < s[i + 1] for i in range(len(s) - 1)):
            raise ValueError(f"invalid sigma interfaces {interfaces!r}")
        self.interfaces = s
        K = self.nlev = len(s) - 1
        self.full_levels = tuple(0.5 * (s[k] + s[k + 1]) for k in range(K))
        self.thickness = tuple(s[k + 1] - s[k] for k in range(K))
        self.log_ratios = (math.inf,) + tuple(math.log(s[k + 1] / s[k]) for k in range(1, K))
        self.alpha = (math.log(2.0),) + tuple(
            1.0 - (s[k] / self.thickness[k]) * self.log_ratios[k] for k in range(1, K))

    @classmethod
    def uniform(cls, nlev: int) -> "Sigma":
        return cls(tuple(k / nlev for k in range(nlev + 1)))


# ---------------------------------------------------------------------------
# Column operators (phy/sigma_coordinate.py)
# ---------------------------------------------------------------------------

def hydrostatic(sg: Sigma, T, phi_s, r_dry: float):
    """(phi_full, phi_below); the upward interface recursion from the surface
    and ``Phi_k = Phi_{k+1/2} + alpha_k R T_k`` (``:164-209``).  Linear in T:
    valid on grid values and on spectral coefficients alike."""
    K = sg.nlev
    phi_below = T * 0.0
    phi_full = T * 0.0
    phi_below[K - 1] = phi_below[K - 1] + phi_s
    for k in range(K - 1, 0, -1):
        phi_below[k - 1] = phi_below[k] + (r_dry * sg.log_ratios[k]) * T[k]
    for k in range(K):
        phi_full[k] = phi_below[k] + (r_dry * sg.alpha[k]) * T[k]
    return phi_full, phi_below


def _weighted(sg: Sigma, G):
    w = G * 0.0
    for k in range(sg.nlev):
        w[k] = sg.thickness[k] * G[k]
    return w


def column_mass_tendency(sg: Sigma, G):
    """``-sum_k Dsigma_k G_k`` (``:221-223``); levels added top to bottom."""
    w = _weighted(sg, G)
    acc = w[0].copy()
    for k in range(1, sg.nlev):
        acc = acc + w[k]
    return -acc


def interface_sigma_dot(sg: Sigma, G):
    """(K+1, ...) interface velocities, rows 0 and K exact zeros (``:226-252``)."""
    w = _weighted(sg, G)
    partial = np.cumsum(w, axis=0)
    total = partial[sg.nlev - 1]
    sd = np.zeros((sg.nlev + 1,) + G.shape[1:], dtype=G.dtype)
    for k in range(1, sg.nlev):
        sd[k] = sg.interfaces[k] * total - partial[k - 1]
    sd[sg.nlev] = total - total
    return sd


def vertical_advection(sg: Sigma, sd, X):
    """Centred ``(sigma_dot dX/dsigma)_k`` (``:280-305``)."""
    K = sg.nlev
    out = X * 0.0
    for k in range(K):
        acc = None
        if k < K - 1:
            acc = sd[k + 1] * (X[k + 1] - X[k])
        if k > 0:
            up = sd[k] * (X[k] - X[k - 1])
            acc = up if acc is None else acc + up
        if acc is not None:
            out[k] = acc / (2.0 * sg.thickness[k])
    return out


def omega_over_p(sg: Sigma, G, A):
    """``A_k - (beta_k/Dsigma_k) P_{k-1} - alpha_k G_k`` (``:373-398``)."""
    w = _weighted(sg, G)
    partial = np.cumsum(w, axis=0)
    out = w * 0.0
    out[0] = A[0] - sg.alpha[0] * G[0]
    for k in range(1, sg.nlev):
        out[k] = (A[k] - (sg.log_ratios[k] / sg.thickness[k]) * partial[k - 1]
                  - sg.alpha[k] * G[k])
    return out


def layer_mass_residual(sg: Sigma, G):
    """``Dsigma_k dlnps/dt + G_k Dsigma_k + (sd_{k+1/2} - sd_{k-1/2})`` (``:462-479``)."""
    w = _weighted(sg, G)
    d = column_mass_tendency(sg, G)
    sd = interface_sigma_dot(sg, G)
    res = w * 0.0
    for k in range(sg.nlev):
        res[k] = sg.thickness[k] * d + w[k] + (sd[k + 1] - sd[k])
    return res


def energy_exchange(sg: Sigma, T, phi_s, G, A, r_dry: float):
    """(conversion, work): ``sum Dsigma R T (omega/p)`` and
    ``sum Dsigma [R T A - (Phi - Phi_s) G]`` (``:401-459``)."""
    phi, _ = hydrostatic(sg, T, phi_s, r_dry)
    wp = omega_over_p(sg, G, A)
    conv = sum((T[k] * wp[k]) * (r_dry * sg.thickness[k]) for k in range(sg.nlev))
    work = sum(sg.thickness[k] * (r_dry * T[k] * A[k] - (phi[k] - phi_s) * G[k])
               for k in range(sg.nlev))
    return conv, work


def vertical_sbp(sg: Sigma, G, X, Y):
    """(lhs, rhs) of the mass-weighted summation-by-parts identity (``:338-370``)."""
    sd = interface_sigma_dot(sg, G)
    ax, ay = vertical_advection(sg, sd, X), vertical_advection(sg, sd, Y)
    d = column_mass_tendency(sg, G)
    lhs = sum(sg.thickness[k] * (X[k] * ay[k] + Y[k] * ax[k]) for k in range(sg.nlev))
    rhs = sum(sg.thickness[k] * (X[k] * Y[k]) * (G[k] + d) for k in range(sg.nlev))
    return lhs, rhs


# ---------------------------------------------------------------------------
# Weak-form vector analysis (num/spectral_operators.py:399-502)
# ---------------------------------------------------------------------------

def extended_transform(planet):
    """The product-sampling transform one degree higher (same points and
    weights; ``num/spectral_operators.py:480-502``)."""
    gp = planet.pgrid
    if isinstance(gp, GaussGrid):
        return GaussTransform(gp, planet.l_max + 1)
    return DenseTransform(gp.latitudes, gp.longitudes, planet.l_max + 1, gp.voronoi_weights,
                          basis="stable")


def c_plus_true(L: int) -> np.ndarray:
    """``C+_{lm} = l sqrt(((l+1)^2 - m^2)/((2l+1)(2l+3)))``, m <= l, without the
    storage zeroing of row L (``num/spectral_operators.py:492-500``)."""
    l = np.arange(L + 1, dtype=np.float64)[:, None] * np.ones((1, L + 1))
    m = np.arange(L + 1, dtype=np.float64)[None, :] * np.ones((L + 1, 1))
    c = l * np.sqrt(np.maximum((l + 1.0) ** 2 - m ** 2, 0.0)
                    / np.maximum((2.0 * l + 1.0) * (2.0 * l + 3.0), 1.0))
    return np.where(m <= l, c, 0.0)


def vector_curl_div(planet, ext, f_east, f_north, truncate: bool = True):
    """``(curl_lm, div_lm)`` of a tangent field on the product sampling
    (``num/spectral_operators.py:399-468``)."""
    so = planet.so
    L = planet.l_max
    n = L + 1
    a = ext.transform(f_east / planet.pcoslat)
    b = ext.transform(f_north / planet.pcoslat)
    cpt = c_plus_true(L)

    def adj(c):
        out = cpt * c[1:n + 1, :n]
        out[1:, :] += so.C_minus[1:, :] * c[:n - 1, :n]
        return out

    div = so.d_lambda(a[:n, :n]) + adj(b) / so.R
    curl = so.d_lambda(b[:n, :n]) - adj(a) / so.R
    if truncate:
        curl, div = so.truncate(curl), so.truncate(div)
    return curl, div


# ---------------------------------------------------------------------------
# Primitive-equation tendency (phy/primitive_equations.py)
# ---------------------------------------------------------------------------

def isothermal_rest_state(L: int, K: int, temperature: float, surface_pressure: float):
    """``phy/primitive_equations.py:139-161``."""
    n = L + 1
    y = np.zeros((3 * K + 1, n, n), dtype=np.complex128)
    mono = math.sqrt(4.0 * math.pi)
    y[2 * K:3 * K, 0, 0] = temperature * mono
    y[3 * K, 0, 0] = math.log(surface_pressure) * mono
    return y


class PEModel:
    """Model constants (``phy/primitive_equations.py:180-264``)."""

    def __init__(self, planet, sigma: Sigma, r_dry: float = R_DRY, cp_dry: float = CP_DRY,
                 phi_surface_lm=None):
        self.planet = planet
        self.sigma = sigma
        self.K = sigma.nlev
        self.r_dry, self.cp_dry = float(r_dry), float(cp_dry)
        self.kappa = self.r_dry / self.cp_dry
        n = planet.l_max + 1
        self.phi_s_lm = (np.zeros((n, n), dtype=np.complex128) if phi_surface_lm is None
                         else np.asarray(phi_surface_lm, dtype=np.complex128))
        self.ext = extended_transform(planet)

    def inv_lap(self, a):
        """l = 0 zeroed (``:266-270``)."""
        lap = self.planet.so.lap_eigs
        out = np.zeros_like(a)
        out[1:, :] = a[1:, :] / lap[1:, None]
        return out

    def derivs(self, a):
        """((1/R) d/dlambda, (1/R) sin(theta) d/dtheta) on the product sampling (``:272-278``)."""
        p, so = self.planet, self.planet.so
        return (p.psh.inv_transform(so.d_lambda(a)),
                p.psh.inv_transform(so.sin_theta_d_theta(a)) / so.R)

    def product_fields(self, y) -> dict:
        """``_tendency_product_fields`` (``:482-587``)."""
        p, sg, K = self.planet, self.sigma, self.K
        c = p.pcoslat
        lnps_lam, lnps_snt = self.derivs(y[3 * K])
        us, vs, zs, ds, ts, advs, gs = [], [], [], [], [], [], []
        for k in range(K):
            psi_lam, psi_snt = self.derivs(self.inv_lap(y[k]))
            chi_lam, chi_snt = self.derivs(self.inv_lap(y[K + k]))
            u = (psi_snt + chi_lam) / c
            v = (psi_lam - chi_snt) / c
            z = p.psh.inv_transform(y[k])
            d = p.psh.inv_transform(y[K + k])
            t = p.psh.inv_transform(y[2 * K + k])
            adv = (u * lnps_lam - v * lnps_snt) / c
            us.append(u); vs.append(v); zs.append(z); ds.append(d); ts.append(t)
            advs.append(adv); gs.append(d + adv)
        u, v, T = np.stack(us), np.stack(vs), np.stack(ts)
        A, G = np.stack(advs), np.stack(gs)
        sd = interface_sigma_dot(sg, G)
        return {"u": u, "v": v, "zeta": np.stack(zs), "delta": np.stack(ds), "temperature": T,
                "lnps_lam": lnps_lam, "lnps_snt": lnps_snt,
                "grad_lnps_u": lnps_lam / c, "grad_lnps_v": -lnps_snt / c,
                "v_grad_lnps": A, "g_full": G, "dlnps_dt": column_mass_tendency(sg, G),
                "sigma_dot": sd, "omega_over_p": omega_over_p(sg, G, A),
                "phi_surface": p.psh.inv_transform(self.phi_s_lm),
                "sigma_dot_dU": vertical_advection(sg, sd, u),
                "sigma_dot_dV": vertical_advection(sg, sd, v),
                "sigma_dot_dT": vertical_advection(sg, sd, T), "coslat": c}

    def thermo_mass(self, y, f):
        """``_thermo_mass_tendencies`` (``:596-635``)."""
        p, K, so = self.planet, self.K, self.planet.so
        c = f["coslat"]
        tds = []
        for k in range(K):
            t_lam, t_snt = self.derivs(y[2 * K + k])
            adv = (f["u"][k] * t_lam - f["v"][k] * t_snt) / c
            g = -adv - f["sigma_dot_dT"][k] + self.kappa * f["temperature"][k] * f["omega_over_p"][k]
            tds.append(so.truncate(p.psh.transform(g)))
        return np.stack(tds), so.truncate(p.psh.transform(f["dlnps_dt"]))

    def momentum(self, y, f):
        """``_momentum_tendencies`` (``:637-725``)."""
        p, K, so = self.planet, self.K, self.planet.so
        c = f["coslat"]
        phi_lm, _ = hydrostatic(self.sigma, y[2 * K:3 * K], self.phi_s_lm, self.r_dry)
        zds, dds = [], []
        for k in range(K):
            eta = y[k] + p.f_lm
            e_lam, e_snt = self.derivs(eta)
            e_g = p.psh.inv_transform(eta)
            u, v = f["u"][k], f["v"][k]
            div_ev = (u * e_lam - v * e_snt) / c + e_g * f["delta"][k]
            curl_ev = e_g * f["zeta"][k] + (v * e_lam + u * e_snt) / c
            rt = self.r_dry * f["temperature"][k]
            ze = f["sigma_dot_dU"][k] + rt * f["grad_lnps_u"]
            zn = f["sigma_dot_dV"][k] + rt * f["grad_lnps_v"]
            curl_z, div_z = vector_curl_div(p, self.ext, ze, zn, truncate=False)
            ke = so.truncate(p.psh.transform(0.5 * (u * u + v * v)))
            zds.append(so.truncate(p.psh.transform(-div_ev) - curl_z))
            dds.append(so.truncate(p.psh.transform(curl_ev) - div_z)
                       - so.lap_eigs[:, None] * (ke + phi_lm[k]))
        zd, dd = np.stack(zds), np.stack(dds)
        zd[:, 0, :] = 0.0
        dd[:, 0, :] = 0.0
        return zd, dd


def state_wind(model: PEModel, y):
    """Per-level (u, v) on the state sampling, clamped ``cos lat``
    (``phy/primitive_equations.py:280-307``)."""
    p, so, K = model.planet, model.planet.so, model.K
    us, vs = [], []
    for k in range(K):
        psi, chi = model.inv_lap(y[k]), model.inv_lap(y[K + k])
        p_lam = p.sh.inv_transform(so.d_lambda(psi))
        p_snt = p.sh.inv_transform(so.sin_theta_d_theta(psi)) / so.R
        c_lam = p.sh.inv_transform(so.d_lambda(chi))
        c_snt = p.sh.inv_transform(so.sin_theta_d_theta(chi)) / so.R
        us.append((p_snt + c_lam) / p.scoslat)
        vs.append((p_lam - c_snt) / p.scoslat)
    return np.stack(us), np.stack(vs)


def temperature_extrema(model: PEModel, y):
    """(min, max) of T over every level on the state and product samplings (``:764-782``)."""
    p, K = model.planet, model.K
    vals = [sh.inv_transform(y[2 * K + k]) for sh in (p.sh, p.psh) for k in range(K)]
    return min(float(v.min()) for v in vals), max(float(v.max()) for v in vals)


def max_characteristic_speed(model: PEModel, y) -> float:
    """``max|V| + sqrt(gamma R_d max(T_max, 0))`` (``:784-806``)."""
    u, v = state_wind(model, y)
    _, t_max = temperature_extrema(model, y)
    gamma = model.cp_dry / (model.cp_dry - model.r_dry)
    return float(np.sqrt(u * u + v * v).max()) + math.sqrt(gamma * model.r_dry * max(t_max, 0.0))


def pe_tendency(model: PEModel, y) -> np.ndarray:
    """``PrimitiveEquationsModel.tendency`` (``phy/primitive_equations.py:727-753``):
    rows ``[zeta_1..K, delta_1..K, T_1..K, ln p_s]``."""
    y = np.asarray(y, dtype=np.complex128)
    f = model.product_fields(y)
    td, ld = model.thermo_mass(y, f)
    zd, dd = model.momentum(y, f)
    return np.concatenate([zd, dd, td, ld[None]], axis=0)

