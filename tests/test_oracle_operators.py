"""Pins of the oracle's spectral operators and Jacobian (CPU).

* S applied to Y_lm equals the analytic sin(theta) dY_lm/dtheta (central
  differences of the stable recurrence) for l < L, and drops the degree-(L+1)
  part at l = L (num/spectral_operators.py:173);
* <S a, b> = <a, S^T b> exactly;
* solid-body J(psi_sb, Y_lm) = i m omega Y_lm, J(a, a) = 0, integral J = 0,
  solid-body velocity (tests/test_backend_parity.py:96-140).
"""
import math

import numpy as np
import pytest

from oracle.basis import normalized_legendre
from oracle.grids import GaussGrid, product_grid_dims
from oracle.models import Planet
from oracle.operators import SpectralOps
from oracle.transforms import GaussTransform

R = 6.371e6


def test_sin_theta_d_theta_matches_analytic_derivative():
    L = 12
    x0 = np.linspace(-0.95, 0.95, 41)
    th = np.arccos(x0)
    h = 1e-6
    Pp = normalized_legendre(np.cos(th + h), L + 1)
    Pm = normalized_legendre(np.cos(th - h), L + 1)
    dP = (Pp - Pm) / (2 * h) * np.sin(th)[None, :]      # sin(theta) d/dtheta, (B_{L+1}, npts)
    so = SpectralOps(L, 1.0)
    P = normalized_legendre(x0, L)
    for l, m in [(3, 1), (5, 0), (7, 7), (11, 4), (12, 2)]:
        a = np.zeros((L + 1, L + 1), dtype=np.complex128)
        a[l, m] = 1.0
        g = so.sin_theta_d_theta(a)
        val = np.zeros_like(x0)
        for ll in range(L + 1):
            val += g[ll, m].real * P[ll * (ll + 1) // 2 + m]
        exact = dP[l * (l + 1) // 2 + m]
        if l < L:
            assert np.abs(val - exact).max() < 1e-8 * max(1.0, np.abs(exact).max())
        else:
            # truncated: the C+ Y_{L+1} part is missing
            Lp = L + 1
            cplus = L * math.sqrt(((L + 1) ** 2 - m * m) / ((2 * L + 1) * (2 * L + 3)))
            missing = cplus * normalized_legendre(x0, Lp)[Lp * (Lp + 1) // 2 + m]
            assert np.abs(val + missing - exact).max() < 1e-8 * max(1.0, np.abs(exact).max())


def test_adjoint_is_transpose():
    L = 17
    so = SpectralOps(L, R)
    rng = np.random.default_rng(0)
    a = rng.standard_normal((L + 1, L + 1)) + 1j * rng.standard_normal((L + 1, L + 1))
    b = rng.standard_normal((L + 1, L + 1)) + 1j * rng.standard_normal((L + 1, L + 1))
    lhs = np.vdot(so.sin_theta_d_theta(a), b)
    rhs = np.vdot(a, so.sin_theta_d_theta_T(b))
    assert abs(lhs - rhs) <= 1e-14 * abs(lhs)


@pytest.fixture(scope="module")
def latlon21():
    return Planet(l_max=21, radius=R, omega=7.292e-5, grid_type="latlon", nlat=32, nlon=64)


@pytest.mark.parametrize("l,m", [(3, 2), (8, 4), (12, 6)])
def test_solid_body_jacobian(latlon21, l, m):
    p = latlon21
    w = 1e-5
    psi = np.zeros((22, 22), dtype=np.complex128)
    psi[1, 0] = -w * R ** 2 * math.sqrt(4 * math.pi / 3)
    q = np.zeros_like(psi)
    q[l, m] = 1.0
    J = p.so.jacobian(psi, q, p.psh, p.pcoslat, dealias=False)
    assert abs(J[l, m] / (1j * m * w) - 1.0) < 1e-8


def test_self_jacobian_and_integral(latlon21):
    p = latlon21
    a = np.zeros((22, 22), dtype=np.complex128)
    a[4, 2] = 1.0 + 0.3j
    a[7, 5] = -0.6 + 0.2j
    assert np.abs(p.so.jacobian(a, a, p.psh, p.pcoslat, dealias=False)).max() < 1e-15
    b = np.zeros_like(a)
    b[3, 1] = 0.7 - 0.2j
    assert abs(p.so.jacobian(a, b, p.psh, p.pcoslat, dealias=False)[0, 0]) < 1e-11


def test_truncation_applied_once_exactly(latlon21):
    p = latlon21
    rng = np.random.default_rng(1)
    a = rng.standard_normal((22, 22)) + 1j * rng.standard_normal((22, 22))
    b = rng.standard_normal((22, 22)) + 1j * rng.standard_normal((22, 22))
    J = p.so.jacobian(a, b, p.psh, p.pcoslat, dealias=True)
    cut = (2 * 21) // 3
    assert np.all(J[cut + 1:, :] == 0) and np.all(J[:, cut + 1:] == 0)


def test_solid_body_velocity(latlon21):
    p = latlon21
    w = 1e-5
    psi = np.zeros((22, 22), dtype=np.complex128)
    psi[1, 0] = -w * R ** 2 * math.sqrt(4 * math.pi / 3)
    u, v = p.so.velocity_from_streamfunction(psi, p.sh, p.grid.point_latitudes)
    lat = p.grid.point_latitudes
    assert np.abs(u - w * R * np.cos(lat)).max() < 1e-8 * w * R
    assert np.abs(v).max() < 1e-8 * w * R


def test_fine_product_quadrature_is_exact_for_products():
    """tests/test_latlon_backend.py:104-178: the 3/2-rule grid analyses the
    product of two degree-L fields exactly (vs a 4x over-resolved grid)."""
    L = 10
    nf, mf = product_grid_dims(L, L + 1, 2 * L + 1)
    fine = GaussTransform(GaussGrid(nf, mf), L)
    big = GaussTransform(GaussGrid(4 * nf, 4 * mf), L)
    rng = np.random.default_rng(3)
    a = rng.standard_normal((L + 1, L + 1)) + 1j * rng.standard_normal((L + 1, L + 1))
    b = rng.standard_normal((L + 1, L + 1)) + 1j * rng.standard_normal((L + 1, L + 1))
    pf = fine.transform(fine.inv_transform(a) * fine.inv_transform(b))
    pb = big.transform(big.inv_transform(a) * big.inv_transform(b))
    assert np.abs(pf - pb).max() < 1e-11 * np.abs(pb).max()
