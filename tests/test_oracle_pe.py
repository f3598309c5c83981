"""Pins of the primitive-equation oracle (CPU): hand values and closed forms of
the sigma-coordinate column operators, the weak-form vector curl/divergence
closed forms and the exact properties of the PE tendency."""
import math

import numpy as np
import pytest

from oracle import pe
from oracle.models import Planet, bve_tendency, planet_constants

NONUNIFORM = pe.Sigma((0.0, 0.05, 0.15, 0.30, 0.50, 0.75, 1.0))
T0, PS0 = 280.0, 1.0e5
SQRT4PI = math.sqrt(4.0 * math.pi)


@pytest.fixture(scope="module")
def latlon():
    R, om = planet_constants(24.0)
    return Planet(l_max=15, radius=R, omega=om, grid_type="latlon", nlat=32, nlon=64)


@pytest.fixture(scope="module")
def geodesic():
    R, om = planet_constants(24.0)
    return Planet(l_max=10, radius=R, omega=om, grid_type="geodesic", resolution=3)


# -- sigma metadata and column operators (tests/test_sigma_coordinate.py) -----

def test_sigma_metadata():
    g = NONUNIFORM
    assert g.alpha[0] == math.log(2.0) and g.log_ratios[0] == math.inf
    for k in range(1, g.nlev):
        assert 0.0 < g.alpha[k] < 1.0
    u = pe.Sigma.uniform(4)
    assert u.interfaces == (0.0, 0.25, 0.5, 0.75, 1.0)
    for bad in [(0.1, 1.0), (0.0, 0.5, 0.999), (0.0, 0.6, 0.4, 1.0), (0.0, 0.5, 0.5, 1.0)]:
        with pytest.raises(ValueError):
            pe.Sigma(bad)


def test_isothermal_geopotential_closed_form():
    """``Phi_{k+1/2} = Phi_s - R T0 ln sigma_{k+1/2}`` and the full level at
    ``ln sigma_{k+1/2} - alpha_k``; the top full level sits at sigma_1
    (``tests/test_sigma_coordinate.py:96-115``)."""
    g, phi_s = NONUNIFORM, 1234.5
    pf, pb = pe.hydrostatic(g, np.full((g.nlev,), 250.0), phi_s, pe.R_DRY)
    assert pb[-1] == phi_s
    for k in range(g.nlev):
        ln_s = math.log(g.interfaces[k + 1])
        assert pb[k] == pytest.approx(phi_s - pe.R_DRY * 250.0 * ln_s, rel=1e-13)
        assert pf[k] == pytest.approx(phi_s - pe.R_DRY * 250.0 * (ln_s - g.alpha[k]), rel=1e-13)
    assert pf[0] == pytest.approx(phi_s - pe.R_DRY * 250.0 * math.log(g.full_levels[0]),
                                  rel=1e-13)


def test_single_layer_sigma_dot_and_uniform_g_omega():
    g = pe.Sigma.uniform(5)
    G = np.zeros((5,))
    G[2] = 1e-5
    w = G[2] * g.thickness[2]
    sd = pe.interface_sigma_dot(g, G)
    for k in range(6):
        assert sd[k] == pytest.approx(g.interfaces[k] * w - (w if k >= 3 else 0.0), abs=1e-21)
    c = 4.0e-6
    wp = pe.omega_over_p(NONUNIFORM, np.full((6, 7), c), np.zeros((6, 7)))
    assert np.allclose(wp[0], -c * math.log(2.0), rtol=1e-14)
    assert np.allclose(wp[1:], -c, rtol=1e-12)
    A = np.random.default_rng(41).standard_normal((5, 6)) * 1e-6
    assert np.array_equal(pe.omega_over_p(g, np.zeros_like(A), A), A)


def test_vertical_advection_hand_values():
    g = pe.Sigma((0.0, 0.3, 1.0))
    sd = np.zeros((3, 2))
    sd[1] = [1.5e-4, -2.0e-4]
    X = np.array([[1.0, 2.0], [4.0, -1.0]])
    out = pe.vertical_advection(g, sd, X)
    np.testing.assert_allclose(out[0], sd[1] * (X[1] - X[0]) / 0.6, rtol=1e-15)
    np.testing.assert_allclose(out[1], sd[1] * (X[1] - X[0]) / 1.4, rtol=1e-15)
    # linear profile on a uniform grid: b times the mean of the two interface velocities
    u = pe.Sigma.uniform(8)
    G = np.random.default_rng(61).standard_normal((8, 6)) * 1e-5
    sdu = pe.interface_sigma_dot(u, G)
    Xl = 1.7 - 4.2 * np.asarray(u.full_levels)[:, None] * np.ones((1, 6))
    out = pe.vertical_advection(u, sdu, Xl)
    for k in range(8):
        np.testing.assert_allclose(out[k], -4.2 * 0.5 * (sdu[k] + sdu[k + 1]), rtol=1e-12,
                                   atol=1e-20)
    assert np.all(pe.vertical_advection(pe.Sigma.uniform(1), np.zeros((2, 3)),
                                        np.ones((1, 3))) == 0.0)


def test_column_closures_are_round_off():
    g = NONUNIFORM
    rng = np.random.default_rng(23)
    G = rng.standard_normal((6, 100)) * 1e-5
    assert np.abs(pe.layer_mass_residual(g, G)).max() < 1e-13 * np.abs(G).max()
    X, Y = rng.standard_normal((6, 100)), rng.standard_normal((6, 100))
    lhs, rhs = pe.vertical_sbp(g, G, X, Y)
    assert np.abs(lhs - rhs).max() < 1e-13 * np.abs(lhs).max()
    T = 220.0 + 60.0 * rng.random((6, 100))
    A = rng.standard_normal((6, 100)) * 1e-6
    conv, work = pe.energy_exchange(g, T, 500.0 * rng.random(100), G, A, pe.R_DRY)
    assert np.abs(conv - work).max() < 1e-12 * np.abs(conv).max()


# -- weak-form vector analysis (tests/test_pe_tendency.py:439-560) ------------

def _potentials(L, modes):
    out = np.zeros((L + 1, L + 1), dtype=np.complex128)
    for l, m, v in modes:
        out[l, m] = v
    return out


def _vector(planet, psi, chi):
    so, psh, c = planet.so, planet.psh, planet.pcoslat
    p_lam, p_snt = psh.inv_transform(so.d_lambda(psi)), psh.inv_transform(
        so.sin_theta_d_theta(psi)) / so.R
    c_lam, c_snt = psh.inv_transform(so.d_lambda(chi)), psh.inv_transform(
        so.sin_theta_d_theta(chi)) / so.R
    return (p_snt + c_lam) / c, (p_lam - c_snt) / c


@pytest.mark.parametrize("psi_modes,chi_modes", [
    ([(1, 0, 4.0e7), (3, 2, 1.0e7 - 5.0e6j), (10, 7, 6.0e6 + 2.0e6j)], []),
    ([], [(2, 0, 3.0e6), (5, 3, -2.0e6 + 1.0e6j), (9, 9, 1.0e6j)]),
    ([(l, m, (1.0 + 0.3 * l) * 1e6 * (1 + 0.5j if m else 1)) for l in range(1, 15)
      for m in range(0, l + 1, 3)],
     [(l, m, (0.7 + 0.1 * m) * 1e6 * (1 - 0.25j if m else 1)) for l in range(1, 15)
      for m in range(1, l + 1, 4)]),
])
def test_weak_form_recovers_curl_and_div(latlon, psi_modes, chi_modes):
    """``curl(k x grad psi + grad chi) = lap psi``, ``div = lap chi`` to round-off
    on the Gauss 3/2-rule product grid for potentials up to l_max - 1 (the
    wind of a degree-l_max potential is already clipped by ``S``)."""
    L = latlon.l_max
    psi, chi = _potentials(L, psi_modes), _potentials(L, chi_modes)
    fe, fn = _vector(latlon, psi, chi)
    ext = pe.extended_transform(latlon)
    curl, div = pe.vector_curl_div(latlon, ext, fe, fn, truncate=False)
    lap = latlon.so.lap_eigs[:, None]
    scale = max(np.abs(lap * psi).max(), np.abs(lap * chi).max())
    assert np.abs(curl - lap * psi).max() < 1e-12 * scale
    assert np.abs(div - lap * chi).max() < 1e-12 * scale
    ct, dt = pe.vector_curl_div(latlon, ext, fe, fn, truncate=True)
    assert np.array_equal(ct, latlon.so.truncate(curl))
    assert np.array_equal(dt, latlon.so.truncate(div))


def test_weak_form_top_degree_is_wind_limited_and_m_confined(latlon):
    """A degree-l_max potential recovers inexactly (its wind lost the
    degree-(l_max+1) part) but only in its own zonal wavenumber, and no
    output has m > l content (``tests/test_pe_tendency.py:496-549``)."""
    L = latlon.l_max
    psi = _potentials(L, [(L, 3, 1.0e7)])
    fe, fn = _vector(latlon, psi, 0 * psi)
    curl, div = pe.vector_curl_div(latlon, pe.extended_transform(latlon), fe, fn,
                                   truncate=False)
    lap = latlon.so.lap_eigs[:, None]
    scale = np.abs(lap * psi).max()
    assert 1e-6 < np.abs(curl - lap * psi).max() / scale < 0.5
    assert 1e-6 < np.abs(div).max() / scale < 0.5
    upper = np.triu(np.ones((L + 1, L + 1)), k=1)
    for out in (curl, div):
        assert np.abs(out).sum() - np.abs(out[:, 3]).sum() < 1e-12 * np.abs(out).max()
        assert not (out * upper).any()


def test_weak_form_zero_field_is_bitwise_zero(latlon):
    z = np.zeros(latlon.pgrid.n_points)
    curl, div = pe.vector_curl_div(latlon, pe.extended_transform(latlon), z, z)
    assert not curl.any() and not div.any()


# -- tendency exact properties (tests/test_pe_tendency.py:793-1000) -----------

def _band_limited(model, seed):
    L, K = model.planet.l_max, model.K
    y = pe.isothermal_rest_state(L, K, T0, PS0)
    rng = np.random.default_rng(seed)
    n, cut = L + 1, (2 * L) // 3

    def pert(amp, zero_mono):
        p = amp * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        p[:, 0] = p[:, 0].real
        p *= np.tril(np.ones((n, n)))
        p[cut + 1:, :] = 0.0
        if zero_mono:
            p[0, 0] = 0.0
        return p
    for k in range(K):
        y[k] += pert(2e-6, True)
        y[K + k] += pert(2e-6, True)
        y[2 * K + k] += pert(0.5, False)
    y[3 * K] += pert(1e-3, True)
    return y


@pytest.mark.parametrize("grid", ["latlon", "geodesic"])
def test_exact_rest_is_bitwise_zero(grid, request):
    planet = request.getfixturevalue(grid)
    m = pe.PEModel(planet, pe.Sigma.uniform(5))
    out = pe.pe_tendency(m, pe.isothermal_rest_state(planet.l_max, 5, T0, PS0))
    assert out.shape == (16, planet.l_max + 1, planet.l_max + 1)
    assert not out.any()


@pytest.mark.parametrize("grid", ["latlon", "geodesic"])
def test_bve_degeneracy(grid, request):
    """delta = 0, per-level uniform T, uniform ln p_s: every level's zeta
    tendency is the barotropic tendency (1e-12 of its scale)."""
    planet = request.getfixturevalue(grid)
    L = planet.l_max
    rng = np.random.default_rng(5)
    z = (rng.standard_normal((L + 1, L + 1)) + 1j * rng.standard_normal((L + 1, L + 1))) * 1e-5
    z = np.tril(z)
    z[:, 0] = z[:, 0].real
    z[0, :] = 0.0
    m = pe.PEModel(planet, pe.Sigma.uniform(4))
    y = pe.isothermal_rest_state(L, 4, T0, PS0)
    for k in range(4):
        y[k] += z
        y[8 + k, 0, 0] = (T0 - 6.0 * k) * SQRT4PI
    out = pe.pe_tendency(m, y)
    ref = bve_tendency(planet, z)
    scale = np.abs(ref).max()
    for k in range(4):
        assert np.abs(out[k] - ref).max() <= 1e-12 * scale


def test_zero_flow_structured_temperature(latlon):
    """Zero flow, uniform ln p_s: zeta tendency bitwise zero, delta tendency
    ``-lap Phi`` of the spectral hydrostatic geopotential."""
    m = pe.PEModel(latlon, pe.Sigma.uniform(5))
    y = pe.isothermal_rest_state(15, 5, T0, PS0)
    for k in range(5):
        y[10 + k, 0, 0] = (T0 - 7.0 * k) * SQRT4PI
        y[10 + k, 4, 2] = 3.0 - 1.0j
        y[10 + k, 2, 0] = 1.5 * (k + 1)
    out = pe.pe_tendency(m, y)
    assert not out[:5].any()
    phi, _ = pe.hydrostatic(m.sigma, y[10:15], m.phi_s_lm, m.r_dry)
    expected = -latlon.so.lap_eigs[None, :, None] * phi
    scale = np.abs(expected).max()
    assert scale > 0 and np.abs(out[5:10] - expected).max() < 1e-12 * scale


def test_monopoles_continuity_and_energy_exchange(latlon):
    m = pe.PEModel(latlon, pe.Sigma.uniform(5))
    y = _band_limited(m, 61)
    out = pe.pe_tendency(m, y)
    assert not out[:10, 0, :].any()
    assert abs(out[15, 0, 0]) > 0 and max(abs(out[10 + k, 0, 0]) for k in range(5)) > 0
    f = m.product_fields(y)
    ref = None
    for k in range(5):
        adv = (f["u"][k] * f["lnps_lam"] - f["v"][k] * f["lnps_snt"]) / f["coslat"]
        term = -m.sigma.thickness[k] * (f["delta"][k] + adv)
        ref = term if ref is None else ref + term
    ref_lm = latlon.so.truncate(latlon.psh.transform(ref))
    assert np.abs(out[15] - ref_lm).max() < 1e-12 * np.abs(ref_lm).max()
    conv, work = pe.energy_exchange(m.sigma, f["temperature"], f["phi_surface"], f["g_full"],
                                    f["v_grad_lnps"], m.r_dry)
    assert np.abs(conv - work).max() < 1e-12 * max(np.abs(conv).max(), np.abs(work).max())


def test_state_wind_and_characteristic_speed(latlon):
    """Solid-body rotation: zeta = 2 w sin(lat) gives u = w R cos(lat), v = 0;
    the resting atmosphere's speed is the Lamb bound sqrt(gamma R_d T0)."""
    m = pe.PEModel(latlon, pe.Sigma.uniform(3))
    y = pe.isothermal_rest_state(15, 3, T0, PS0)
    w = 1e-5
    y[1, 1, 0] = 2.0 * w * math.sqrt(4.0 * math.pi / 3.0)
    u, v = pe.state_wind(m, y)
    lat = latlon.grid.point_latitudes
    np.testing.assert_allclose(u[1], w * latlon.R * np.cos(lat), rtol=1e-10, atol=1e-9)
    assert np.abs(v).max() < 1e-9 and np.abs(u[0]).max() == 0.0
    gamma = pe.CP_DRY / (pe.CP_DRY - pe.R_DRY)
    rest = pe.isothermal_rest_state(15, 3, T0, PS0)
    lamb = math.sqrt(gamma * pe.R_DRY * T0)
    assert pe.max_characteristic_speed(m, rest) == pytest.approx(lamb, rel=1e-9)
    assert pe.max_characteristic_speed(m, y) == pytest.approx(
        w * latlon.R * np.cos(lat).max() + lamb, rel=1e-9)


def test_pe_record_row_resting_atmosphere(latlon):
    """Uniform p_s: total mass 4 pi R^2 p_s, zero drift and wind, Lamb speed."""
    from oracle.diagnostics import pe_record_row
    m = pe.PEModel(latlon, pe.Sigma.uniform(4))
    y = pe.isothermal_rest_state(15, 4, T0, PS0)
    row = pe_record_row(m, y, 300.0, 1.0e5)
    assert row["total_mass"] == pytest.approx(4.0 * math.pi * latlon.R ** 2 * PS0, rel=1e-13)
    assert row["mass_rel_drift"] == 0.0 and row["max_wind_ms"] == 0.0
    assert row["ps_min"] == pytest.approx(PS0, rel=1e-13)
    assert row["t_min"] == pytest.approx(T0, rel=1e-13)
    lamb = math.sqrt(pe.CP_DRY / (pe.CP_DRY - pe.R_DRY) * pe.R_DRY * T0)
    assert row["max_char_speed_ms"] == pytest.approx(lamb, rel=1e-12)
    assert row["courant"] == pytest.approx(lamb * 300.0 / 1.0e5, rel=1e-12)
