"""Spectral invariants (oracle, test infrastructure only).

Follows ``run/bve/diagnostics.py:52-97``: per-mode power
``P = Re(a_l0)^2`` for m = 0 and ``2 |a_lm|^2`` for m > 0, padding (m > l)
ignored; ``E = 1/2 R^4 sum_l P_l/(l(l+1))``; ``Z = 1/2 R^2 sum_l P_l``;
absolute enstrophy with ``f`` added to the (1, 0) coefficient;
circulation ``R^2 sqrt(4 pi) Re a_00``.  Sums use ``math.fsum`` (exactly
rounded), the yardstick for the device's compensated sums.

Pinned by (tests/test_oracle_models.py): the closed forms of
``tests/test_diagnostics.py:63-108`` (single mode, circulation, absolute
enstrophy of f, padding ignored).

``bve_record_row`` / ``swe_record_row`` restate one row of
``DiagnosticsRecorder.record`` (``run/bve/diagnostics.py:162-211``) and
``SWEDiagnosticsRecorder.record`` (``run/swe/diagnostics.py:113-173``);
pinned (tests/test_oracle_diagnostics.py) by solid-body rotation (speed,
energy, rms vorticity, CFL in closed form) and the resting layer (mass,
energy, depth, free surface).

``pe_record_row`` restates one row of ``PEDiagnosticsRecorder.record``
(``run/pe/diagnostics.py:86-129``); pinned (tests/test_oracle_pe.py) by the
resting atmosphere (uniform p_s: mass ``4 pi R^2 p_s`` and zero drift, zero
wind, the Lamb-bound speed and its Courant number).
"""
from __future__ import annotations

import math

import numpy as np

__all__ = ["mode_power", "spectral_diagnostics", "bve_record_row", "swe_record_row",
           "pe_record_row"]


def mode_power(a: np.ndarray) -> np.ndarray:
    n_l, n_m = a.shape
    li, mi = np.indices((n_l, n_m))
    p = np.where(mi == 0, a.real ** 2, 2.0 * np.abs(a) ** 2)
    return np.where(mi <= li, p, 0.0)


def spectral_diagnostics(zeta: np.ndarray, radius: float, omega: float) -> dict:
    R = float(radius)
    L = zeta.shape[0] - 1
    P = mode_power(zeta)
    Pl = np.array([math.fsum(row) for row in P])
    l = np.arange(L + 1, dtype=np.float64)
    inv = np.zeros(L + 1)
    inv[1:] = 1.0 / (l[1:] * (l[1:] + 1.0))
    q = zeta.copy()
    q[1, 0] = q[1, 0] + 2.0 * omega * math.sqrt(4.0 * math.pi / 3.0)
    return {
        "circulation": R ** 2 * math.sqrt(4.0 * math.pi) * float(zeta[0, 0].real),
        "energy": 0.5 * R ** 4 * math.fsum(inv * Pl),
        "enstrophy_rel": 0.5 * R ** 2 * math.fsum(Pl),
        "enstrophy_abs": 0.5 * R ** 2 * math.fsum(mode_power(q).ravel()),
    }


def bve_record_row(planet, zeta: np.ndarray, dt: float, length_scale) -> dict:
    """One ``DiagnosticsRecorder.record`` row with the round-trip residual
    (``run/bve/diagnostics.py:162-211``)."""
    R, L = planet.R, planet.l_max
    spec = spectral_diagnostics(zeta, R, planet.Omega)
    P = mode_power(zeta)
    Pl = np.array([math.fsum(row) for row in P])
    l = np.arange(L + 1, dtype=np.float64)
    Z_l = 0.5 * R ** 2 * Pl
    E_l = np.zeros(L + 1)
    E_l[1:] = 0.5 * R ** 4 * Pl[1:] / (l[1:] * (l[1:] + 1.0))
    cut = (2 * L) // 3
    sh, so = planet.sh, planet.so
    z = sh.inv_transform(zeta)
    w = np.asarray(sh.weights, dtype=np.float64).ravel()
    psi = so.inv_laplacian(zeta)
    u, v = so.velocity_from_streamfunction(psi, sh, planet.grid.point_latitudes)
    max_speed = float(np.max(np.sqrt(u * u + v * v)))
    back = sh.transform(z)
    return {
        "max_speed_ms": max_speed,
        "cfl": max_speed * dt / length_scale if (length_scale and dt > 0) else float("nan"),
        "circulation": spec["circulation"],
        "energy": spec["energy"],
        "enstrophy_rel": spec["enstrophy_rel"],
        "enstrophy_abs": spec["enstrophy_abs"],
        "zeta_max": float(np.max(np.abs(z))),
        "zeta_rms": math.sqrt(math.fsum(w * z * z) / math.fsum(w)),
        "energy_l1": float(E_l[1]) if L >= 1 else 0.0,
        "high_l_enstrophy_frac": (math.fsum(Z_l[cut + 1:]) / spec["enstrophy_rel"]
                                  if spec["enstrophy_rel"] > 0 else 0.0),
        "roundtrip_residual": float(np.linalg.norm(back - zeta)) / (float(np.linalg.norm(zeta))
                                                                    + 1e-300),
    }


def swe_record_row(p, y: np.ndarray, dt: float, length_scale, terrain_max_m: float = 0.0) -> dict:
    """One ``SWEDiagnosticsRecorder.record`` row (``run/swe/diagnostics.py:113-173``,
    ``characteristic_fields`` ``phy/shallow_water.py:456-468``); ``p`` is a
    ``SWEParams``, ``y`` the (3, L+1, L+1) stack."""
    pl = p.planet
    R, so, sh = pl.R, pl.so, pl.sh
    zeta, delta, phi = y[0], y[1], y[2]
    psi, chi = p.inv_lap(zeta), p.inv_lap(delta)
    c = pl.scoslat

    def deriv(a):
        return sh.inv_transform(so.d_lambda(a)), sh.inv_transform(so.sin_theta_d_theta(a)) / R

    psi_lam, psi_snt = deriv(psi)
    chi_lam, chi_snt = deriv(chi)
    u = (psi_snt + chi_lam) / c
    v = (psi_lam - chi_snt) / c
    phi_total = p.phi0 + sh.inv_transform(phi)
    lo, hi = float(phi_total.min()), float(phi_total.max())
    if pl.psh is not pl.sh:
        pt = p.phi0 + pl.psh.inv_transform(phi)
        lo, hi = min(lo, float(pt.min())), max(hi, float(pt.max()))
    max_wind = float(np.max(np.sqrt(u * u + v * v)))
    max_char = max_wind + math.sqrt(max(hi, 0.0))
    kinetic = 0.5 * phi_total * (u * u + v * v)
    potential = 0.5 * phi_total ** 2
    if p.phi_s_lm is not None:
        phi_s = sh.inv_transform(p.phi_s_lm)
        mean_phi_s = float(p.phi_s_lm[0, 0].real) / math.sqrt(4.0 * math.pi)
        potential = potential + phi_total * phi_s
        eta = (phi_total - p.phi0 + (phi_s - mean_phi_s)) / p.gravity
    else:
        eta = (phi_total - p.phi0) / p.gravity
    w = np.asarray(sh.weights, dtype=np.float64).ravel() * R ** 2
    return {
        "max_wind_ms": max_wind,
        "max_char_speed_ms": max_char,
        "cfl": max_char * dt / length_scale if (length_scale and dt > 0) else float("nan"),
        "phi_total_min": lo,
        "phi_total_max": hi,
        "total_mass": R ** 2 * (4.0 * math.pi * p.phi0
                                + math.sqrt(4.0 * math.pi) * float(phi[0, 0].real)),
        "total_energy": math.fsum(w * (kinetic + potential)),
        "h_min_m": lo / p.gravity,
        "eta_min_m": float(eta.min()),
        "eta_max_m": float(eta.max()),
        "terrain_max_m": float(terrain_max_m),
        "zeta_l2": math.sqrt(R ** 2 * math.fsum(mode_power(zeta).ravel())),
        "delta_l2": math.sqrt(R ** 2 * math.fsum(mode_power(delta).ravel())),
        "phi_l2": math.sqrt(R ** 2 * math.fsum(mode_power(phi).ravel())),
    }


def pe_record_row(model, y: np.ndarray, dt: float, length_scale, mass0=None) -> dict:
    """One PE diagnostics row (``run/pe/diagnostics.py:86-129``); ``model`` is an
    ``oracle.pe.PEModel``, ``mass0`` the first row's total mass (None: this row)."""
    from . import pe
    p, K = model.planet, model.K
    t_min, t_max = pe.temperature_extrema(model, y)
    ps = np.exp(p.sh.inv_transform(y[3 * K]))
    u, v = pe.state_wind(model, y)
    zeta = np.stack([p.sh.inv_transform(y[k]) for k in range(K)])
    delta = np.stack([p.sh.inv_transform(y[K + k]) for k in range(K)])
    max_char = pe.max_characteristic_speed(model, y)
    courant = max_char * dt / length_scale if (length_scale and dt > 0) else float("nan")
    mass = math.fsum(p.sh.weights * p.R ** 2 * ps)
    m0 = mass if mass0 is None else mass0
    return {"t_min": t_min, "t_max": t_max, "ps_min": float(ps.min()), "ps_max": float(ps.max()),
            "max_wind_ms": float(np.sqrt(u * u + v * v).max()),
            "max_abs_zeta": float(np.abs(zeta).max()), "max_abs_delta": float(np.abs(delta).max()),
            "max_char_speed_ms": max_char, "courant": courant, "total_mass": mass,
            "mass_rel_drift": (mass - m0) / m0 if m0 else 0.0}
