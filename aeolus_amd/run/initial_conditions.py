"""Seeded initial conditions for the fixed-dt workloads (host NumPy, init only).

* :func:`spectrum_l15` draws a real-field coefficient array with
  ``|a_lm| = A l^-1.5`` for 1 <= l <= L, uniform phases, real m = 0 entries
  ``A l^-1.5 cos(phase)``, zero l = 0, scaled to a target rms
  ``sqrt(sum_l P_l / 4 pi)``.
* :func:`gaussian_vortex_pair` is two opposite-signed e-folding Gaussian
  vortices with seeded centres (the sign pattern of the reference's
  ``two_vortices``, ``run/bve/initial_conditions.py:29-34``).
* :func:`rh4` is the Rossby-Haurwitz wave 4 vorticity
  (``run/bve/initial_conditions.py:75-97``).
* :func:`swe_spectra` builds the shallow-water (zeta, delta, phi') stack with
  rms wind 20 m/s, rms(delta) = 0.1 rms(zeta) and a 100 m thickness spectrum.
"""
from __future__ import annotations

import math

import numpy as np

__all__ = ["power_per_degree", "field_rms", "spectrum_l15", "gaussian_vortex_pair", "rh4",
           "swe_spectra", "random_unit_centre"]


def power_per_degree(a: np.ndarray) -> np.ndarray:
    L = a.shape[-1] - 1
    li, mi = np.indices((L + 1, L + 1))
    p = np.where(mi == 0, a.real ** 2, 2.0 * np.abs(a) ** 2)
    return np.where(mi <= li, p, 0.0).sum(axis=-1)


def field_rms(a: np.ndarray) -> float:
    return math.sqrt(float(power_per_degree(a).sum()) / (4.0 * math.pi))


def spectrum_l15(l_max: int, rng: np.random.Generator, rms: float | None = None) -> np.ndarray:
    L = int(l_max)
    phases = rng.uniform(0.0, 2.0 * math.pi, size=(L + 1, L + 1))
    l = np.arange(L + 1, dtype=np.float64)
    amp = np.zeros(L + 1)
    amp[1:] = l[1:] ** -1.5
    a = amp[:, None] * np.exp(1j * phases)
    a[:, 0] = amp * np.cos(phases[:, 0])
    li, mi = np.indices((L + 1, L + 1))
    a = np.where(mi <= li, a, 0.0).astype(np.complex128)
    a[0, :] = 0.0
    if rms is not None:
        a *= rms / field_rms(a)
    return a


def random_unit_centre(rng: np.random.Generator) -> tuple[float, float]:
    """(lat, lon) uniform on the sphere."""
    z = rng.uniform(-1.0, 1.0)
    lon = rng.uniform(0.0, 2.0 * math.pi)
    return math.asin(z), lon


def gaussian_vortex_pair(lat: np.ndarray, lon: np.ndarray, radius: float, seed: int,
                         peak: float = 1e-4, efold_m: float = 6e5) -> np.ndarray:
    rng = np.random.default_rng(seed)
    (la1, lo1), (la2, lo2) = random_unit_centre(rng), random_unit_centre(rng)

    def bump(la0, lo0):
        cosd = np.sin(lat) * math.sin(la0) + np.cos(lat) * math.cos(la0) * np.cos(lon - lo0)
        d = np.arccos(np.clip(cosd, -1.0, 1.0))
        return np.exp(-(radius * d / efold_m) ** 2)

    return peak * bump(la1, lo1) - peak * bump(la2, lo2)


def rh4(lat: np.ndarray, lon: np.ndarray) -> np.ndarray:
    nu = K = 7.848e-6
    s, c = np.sin(lat), np.cos(lat)
    return 2.0 * nu * s - 30.0 * K * s * c ** 4 * np.sin(4.0 * lon)


def swe_spectra(l_max: int, radius: float, gravity: float, seed: int, *, u_rms: float = 20.0,
                delta_ratio: float = 0.1, h_rms: float = 100.0) -> np.ndarray:
    """(3, L+1, L+1) stack: zeta, delta with rms(u) = u_rms; phi' = g h'."""
    rng = np.random.default_rng(seed)
    zeta = spectrum_l15(l_max, rng, rms=1.0)
    delta = spectrum_l15(l_max, rng, rms=delta_ratio)
    h = spectrum_l15(l_max, rng, rms=h_rms)
    l = np.arange(l_max + 1, dtype=np.float64)
    w = np.zeros(l_max + 1)
    w[1:] = radius ** 2 / (l[1:] * (l[1:] + 1.0))
    u2 = float(np.sum(w * (power_per_degree(zeta) + power_per_degree(delta)))) / (4.0 * math.pi)
    s = u_rms / math.sqrt(u2)
    return np.stack([zeta * s, delta * s, gravity * h])


# ---------------------------------------------------------------------------
# Primitive-equation presets (run/pe/initial_conditions.py:57-131): built on
# the exact isothermal rest state in spectral space (no grid round trip).
# ---------------------------------------------------------------------------
THERMAL_WAVE_DEGREE = 2
THERMAL_WAVE_ORDER = 2


def _pe_isothermal_rest(model, *, temperature, surface_pressure, thermal_amplitude=0.0):
    from ..physics.primitive_equations import isothermal_rest_state
    return isothermal_rest_state(model.l_max, model.nlev, temperature=temperature,
                                 surface_pressure=surface_pressure, device=model.device)


def _pe_thermal_wave(model, *, temperature, surface_pressure, thermal_amplitude):
    """Rest plus one real coefficient ``thermal_amplitude`` (K) on the (2, 2)
    mode at every level: unbalanced by design."""
    if model.l_max < THERMAL_WAVE_DEGREE:
        raise ValueError(f"thermal_wave needs l_max >= {THERMAL_WAVE_DEGREE}, got {model.l_max}")
    st = _pe_isothermal_rest(model, temperature=temperature, surface_pressure=surface_pressure)
    st.temperature[:, THERMAL_WAVE_DEGREE, THERMAL_WAVE_ORDER] = float(thermal_amplitude)
    return st


def _pe_orographic_isothermal_rest(model, *, temperature, surface_pressure,
                                   thermal_amplitude=0.0):
    """Rest balanced over the model's surface geopotential:
    ``ln p_s = ln p_ref - Phi_s / (R_d T0)`` on the spectral coefficients."""
    st = _pe_isothermal_rest(model, temperature=temperature, surface_pressure=surface_pressure)
    st.coeffs[3 * model.nlev] -= model.phi_surface_lm / (model.r_dry * float(temperature))
    return st


PE_INITIAL_CONDITIONS = {
    "isothermal_rest": _pe_isothermal_rest,
    "thermal_wave": _pe_thermal_wave,
    "orographic_isothermal_rest": _pe_orographic_isothermal_rest,
}


def make_pe_ic(name: str, model, *, temperature: float, surface_pressure: float,
               thermal_amplitude: float = 0.0):
    """A named primitive-equation initial state (``run/pe/initial_conditions.py:115-131``)."""
    if name not in PE_INITIAL_CONDITIONS:
        raise ValueError(f"Unknown pe initial condition: {name}. "
                         f"Available: {sorted(PE_INITIAL_CONDITIONS)}")
    return PE_INITIAL_CONDITIONS[name](model, temperature=temperature,
                                       surface_pressure=surface_pressure,
                                       thermal_amplitude=thermal_amplitude)
