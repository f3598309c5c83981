"""Pins of the oracle's tangent-linear and adjoint BVE step (CPU).

The adjoint has no reference counterpart; it is pinned by what the
mathematics fixes: dot-product identities per operator and for the whole
step, the exact polarisation identity of the quadratic tendency, central
finite differences, and planted mistakes that the identities must reject.
"""
import math

import numpy as np
import pytest

from oracle import adjoint as adj
from oracle.models import Planet, bve_tendency, rk4_step

R = 6.371e6
OMEGA = 2.0 * math.pi / 86400.0


def _field(L, rng, rms=5e-5, valid=True):
    """Seeded l^-1.5 field; valid=False fills every entry (m > l, Im a_l0 too)."""
    l = np.arange(L + 1, dtype=np.float64)
    amp = np.zeros(L + 1)
    amp[1:] = l[1:] ** -1.5
    a = amp[:, None] * np.exp(1j * rng.uniform(0, 2 * math.pi, (L + 1, L + 1)))
    if valid:
        a[:, 0] = a[:, 0].real
        li, mi = np.indices(a.shape)
        a = np.where(mi <= li, a, 0.0)
        a[0, :] = 0.0
    else:
        a = a + rng.standard_normal(a.shape) * 1e-2 + 1j * rng.standard_normal(a.shape) * 1e-2
    return a * (rms / np.sqrt(np.sum(np.abs(a) ** 2)))


@pytest.fixture(scope="module", params=["latlon", "geodesic"])
def planet(request):
    if request.param == "latlon":
        return Planet(l_max=10, radius=R, omega=OMEGA, grid_type="latlon", nlat=16, nlon=32)
    return Planet(l_max=6, radius=R, omega=OMEGA, grid_type="geodesic", resolution=2)


def _rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def test_transform_transposes(planet):
    rng = np.random.default_rng(1)
    for sh in (planet.sh, planet.psh):
        L = sh.l_max
        a = _field(L, rng, valid=False)
        g = rng.standard_normal(sh.n_points)
        assert _rel(adj.inner(sh.inv_transform(a), g), adj.inner(a, adj.synthesis_T(sh, g))) < 1e-13
        f = rng.standard_normal(sh.n_points)
        ab = _field(L, rng, valid=False)
        assert _rel(adj.inner(sh.transform(f), ab), adj.inner(f, adj.analysis_T(sh, ab))) < 1e-13


def test_spectral_operator_transposes(planet):
    rng = np.random.default_rng(2)
    so = planet.so
    a, b = _field(so.l_max, rng, valid=False), _field(so.l_max, rng, valid=False)
    assert _rel(adj.inner(so.d_lambda(a), b), adj.inner(a, adj.d_lambda_T(so, b))) < 1e-14
    assert _rel(adj.inner(so.sin_theta_d_theta(a), b),
                adj.inner(a, so.sin_theta_d_theta_T(b))) < 1e-14


def test_jacobian_tangent_and_transpose(planet):
    rng = np.random.default_rng(3)
    so, psh, c = planet.so, planet.psh, planet.pcoslat
    L = so.l_max
    a, b, da, db = (_field(L, rng) for _ in range(4))
    # bilinear: the tangent is exact, J(a+da, b+db) - J(a-da, b-db) = 2 J'(da, db)
    t = adj.jacobian_tl(so, a, b, da, db, psh, c)
    pol = 0.5 * (so.jacobian(a + da, b + db, psh, c) - so.jacobian(a - da, b - db, psh, c))
    assert np.linalg.norm(t - pol) <= 1e-12 * np.linalg.norm(t)
    jb = _field(L, rng, valid=False)
    abar, bbar = adj.jacobian_T(so, a, b, jb, psh, c)
    lhs = adj.inner(t, jb)
    rhs = adj.inner(da, abar) + adj.inner(db, bbar)
    assert _rel(lhs, rhs) < 1e-12


@pytest.mark.parametrize("nu", [0.0, 1e5])
def test_tendency_tangent_and_adjoint(planet, nu):
    rng = np.random.default_rng(4)
    L = planet.l_max
    z, d = _field(L, rng), _field(L, rng)
    T = lambda y: bve_tendency(planet, y, nu)
    tl = adj.bve_tendency_tl(planet, z, d, nu)
    # T is quadratic + linear: (T(z+d) - T(z-d))/2 = T'(z) d exactly
    pol = 0.5 * (T(z + d) - T(z - d))
    assert np.linalg.norm(tl - pol) <= 1e-12 * np.linalg.norm(tl)
    eps = 1e-4
    fd = (T(z + eps * d) - T(z - eps * d)) / (2 * eps)
    assert np.linalg.norm(tl - fd) <= 1e-9 * np.linalg.norm(tl)
    kb = _field(L, rng, valid=False)
    zb = adj.bve_tendency_T(planet, z, kb, nu)
    assert _rel(adj.inner(tl, kb), adj.inner(d, zb)) < 1e-12


def test_rk4_step_dot_product_and_gradient(planet):
    rng = np.random.default_rng(5)
    L, nu, dt = planet.l_max, 2e5, 900.0
    T, Ttl, TT = adj.tendency_funcs(planet, nu)
    z, dz = _field(L, rng), _field(L, rng)
    yb = _field(L, rng, valid=False)
    y1, stages = adj.rk4_step_record(T, z, dt)
    assert np.array_equal(y1, rk4_step(T, z, dt))
    _, dy = adj.rk4_step_tl(T, Ttl, z, dz, dt)
    g = adj.rk4_step_T(TT, stages, yb, dt)
    assert _rel(adj.inner(dy, yb), adj.inner(dz, g)) < 1e-12
    # misfit gradient of a 3-step run against central differences
    target = _field(L, np.random.default_rng(6))

    def run(z0):
        y, st = z0, []
        for _ in range(3):
            y, s = adj.rk4_step_record(T, y, dt)
            st.append(s)
        return y, st

    yT, st = run(z)
    J, gbar = adj.sphere_misfit(yT, target)
    for s in reversed(st):
        gbar = adj.rk4_step_T(TT, s, gbar, dt)
    eps = 1e-7 * np.linalg.norm(z)
    for k in range(4):
        d = _field(L, np.random.default_rng(100 + k))
        d /= np.linalg.norm(d)
        jp = adj.sphere_misfit(run(z + eps * d)[0], target)[0]
        jm = adj.sphere_misfit(run(z - eps * d)[0], target)[0]
        fd = (jp - jm) / (2 * eps)
        assert _rel(fd, adj.inner(gbar, d)) < 1e-6


def test_sphere_misfit_is_the_l2_norm_of_the_field():
    # 1/2 int (f - g)^2 dOmega on an exact Gauss quadrature equals the misfit
    p = Planet(l_max=8, radius=R, omega=OMEGA, grid_type="latlon", nlat=12, nlon=24)
    rng = np.random.default_rng(7)
    a, b = _field(8, rng), _field(8, rng)
    J, _ = adj.sphere_misfit(a, b)
    d = p.sh.inv_transform(a) - p.sh.inv_transform(b)
    assert _rel(J, 0.5 * float(np.sum(p.sh.weights * d * d))) < 1e-12


def test_dot_product_rejects_planted_mistakes(planet, monkeypatch):
    """A dropped c_m factor or S in place of S^T must break the identity."""
    rng = np.random.default_rng(8)
    L = planet.l_max
    z, d, kb = _field(L, rng), _field(L, rng), _field(L, rng)
    tl = adj.bve_tendency_tl(planet, z, d)

    def mismatch():
        return _rel(adj.inner(tl, kb), adj.inner(d, adj.bve_tendency_T(planet, z, kb)))

    assert mismatch() < 1e-12
    with monkeypatch.context() as mp:
        mp.setattr(adj, "synthesis_T", _synthesis_T_without_cm)
        assert mismatch() > 1e-3
    so = planet.so
    monkeypatch.setattr(so, "sin_theta_d_theta_T", so.sin_theta_d_theta)
    assert mismatch() > 1e-3


def _synthesis_T_without_cm(sh, g):
    g = np.asarray(g, dtype=np.float64).ravel()
    return sh.transform(g / sh.weights)
