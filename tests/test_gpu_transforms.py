"""HIP transforms, spectral operators and Jacobian against the CPU oracle.

Bars (written in each assertion): FP64 parity to the oracle within 1e-13
relative L2 for Gauss transforms (same sums, different order), 1e-12 for
the dense point-set path; the reference known answers of
tests/test_latlon_transform.py:99-121 and tests/test_backend_parity.py:81-140.
"""
import math

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

GAUSS_CASES = [(12, 24, 8), (13, 25, 8), (24, 48, 10), (32, 64, 21), (64, 128, 42),
               (128, 256, 85), (256, 512, 170), (31, 47, 15)]


def _rand_coeffs(L, rng, amp=1.0):
    a = np.zeros((L + 1, L + 1), dtype=np.complex128)
    for l in range(L + 1):
        a[l, 0] = amp * rng.standard_normal()
        for m in range(1, l + 1):
            a[l, m] = amp * (rng.standard_normal() + 1j * rng.standard_normal())
    return a


def _np(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    return torch


@pytest.mark.parametrize("nlat,nlon,L", GAUSS_CASES)
def test_gauss_synthesis_and_analysis_match_oracle(torch_mod, nlat, nlon, L):
    from aeolus_amd.numerics import GaussLatLonGridGeometry, GaussLatLonSphericalHarmonics
    from oracle.grids import GaussGrid
    from oracle.transforms import GaussTransform
    print(f"case gauss {nlat}x{nlon} L={L}", flush=True)
    rng = np.random.default_rng(nlat * 1000 + L)
    g = GaussLatLonGridGeometry(nlat, nlon, 1.0)
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sh = GaussLatLonSphericalHarmonics(g, L)
    og = GaussGrid(nlat, nlon)
    ot = GaussTransform(og, L)
    a = _rand_coeffs(L, rng)
    f_gpu = _np(sh.inv_transform(torch_mod.as_tensor(a, device="cuda")))
    f_ref = ot.inv_transform(a)
    assert rel_l2(f_gpu, f_ref) < 1e-13
    f = rng.standard_normal(nlat * nlon)
    a_gpu = _np(sh.transform(torch_mod.as_tensor(f, device="cuda")))
    a_ref = ot.transform(f)
    assert rel_l2(a_gpu, a_ref) < 1e-13
    li, mi = np.indices((L + 1, L + 1))
    assert np.all(a_gpu[mi > li] == 0)


def test_gauss_known_answers(torch_mod):
    """tests/test_latlon_transform.py:99-121 on the HIP transform."""
    from aeolus_amd.numerics import GaussLatLonGridGeometry, GaussLatLonSphericalHarmonics
    L = 10
    g = GaussLatLonGridGeometry(24, 48, 6.371e6)
    sh = GaussLatLonSphericalHarmonics(g, L)
    rng = np.random.default_rng(11)
    a = _rand_coeffs(L, rng, 1e-5)
    back = _np(sh.transform(sh.inv_transform(torch_mod.as_tensor(a, device="cuda"))))
    assert np.abs(back - a).max() / np.abs(a).max() < 1e-10
    d = np.zeros((L + 1, L + 1), dtype=np.complex128)
    d[3, 2] = 1.0 - 0.5j
    back = _np(sh.transform(sh.inv_transform(torch_mod.as_tensor(d, device="cuda"))))
    assert abs(back[3, 2] - (1.0 - 0.5j)) < 1e-11
    back[3, 2] = 0
    assert np.abs(back).max() < 1e-11
    ones = torch_mod.ones(g.n_points, dtype=torch_mod.float64, device="cuda")
    a00 = complex(_np(sh.transform(ones))[0, 0])
    assert abs(a00 - math.sqrt(4 * math.pi)) / math.sqrt(4 * math.pi) < 1e-13


@pytest.mark.parametrize("res,L", [(2, 5), (3, 10), (4, 21)])
def test_points_transform_matches_oracle(torch_mod, res, L):
    from aeolus_amd.numerics import GeodesicGridGeometry, GeodesicSphericalHarmonics
    from oracle.grids import GeodesicGrid
    from oracle.transforms import DenseTransform
    print(f"case points res={res} L={L}", flush=True)
    import warnings
    g = GeodesicGridGeometry(res, 1.0)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        sh = GeodesicSphericalHarmonics(g, L)
    og = GeodesicGrid(res, 1.0)
    # SciPy's rotation and the oracle's closed-form matrix agree to rounding.
    assert np.abs(g._latitudes_np - og.latitudes).max() < 1e-14
    assert np.abs(np.exp(1j * g._longitudes_np) - np.exp(1j * og.longitudes)).max() < 1e-14
    for basis in ("reference", "stable"):
        ot = DenseTransform(og.latitudes, og.longitudes, L, og.voronoi_weights, basis=basis)
        rng = np.random.default_rng(res)
        a = _rand_coeffs(L, rng)
        f_gpu = _np(sh.inv_transform(torch_mod.as_tensor(a, device="cuda")))
        assert rel_l2(f_gpu, ot.inv_transform(a)) < 1e-12
        f = rng.standard_normal(og.n_points)
        a_gpu = _np(sh.transform(torch_mod.as_tensor(f, device="cuda")))
        assert rel_l2(a_gpu, ot.transform(f)) < 1e-12
    Y = _np(sh.Y_matrix)
    ot = DenseTransform(og.latitudes, og.longitudes, L, og.voronoi_weights, basis="reference")
    assert np.abs(Y - ot.Y).max() < 1e-12


@pytest.mark.parametrize("op", ["laplacian_coeffs", "inv_laplacian", "d_lambda_coeffs",
                                "sin_theta_d_theta_coeffs", "adjoint_sin_theta_d_theta_coeffs",
                                "_truncate_product"])
def test_spectral_ops_match_oracle(torch_mod, op):
    from aeolus_amd.numerics import (GaussLatLonGridGeometry, GaussLatLonSphericalHarmonics,
                                     SpectralOperators)
    from oracle.operators import SpectralOps
    L, R = 21, 6.371e6
    g = GaussLatLonGridGeometry(32, 64, R)
    sh = GaussLatLonSphericalHarmonics(g, L)
    so = SpectralOperators(sh, R, g, product_quadrature="fine")
    oo = SpectralOps(L, R)
    a = _rand_coeffs(L, np.random.default_rng(5))
    got = _np(getattr(so, op)(torch_mod.as_tensor(a, device="cuda")))
    ref = {"laplacian_coeffs": oo.laplacian, "inv_laplacian": oo.inv_laplacian,
           "d_lambda_coeffs": oo.d_lambda, "sin_theta_d_theta_coeffs": oo.sin_theta_d_theta,
           "adjoint_sin_theta_d_theta_coeffs": oo.sin_theta_d_theta_T,
           "_truncate_product": oo.truncate}[op](a)
    assert rel_l2(got, ref) < 1e-15


@pytest.mark.parametrize("kind", ["latlon", "geodesic"])
def test_jacobian_matches_oracle(torch_mod, kind):
    from aeolus_amd.planet import Planet
    from oracle.models import Planet as OPlanet
    L, R, Om = 21, 6.371e6, 7.292e-5
    kw = dict(grid_type=kind, nlat=32, nlon=64, resolution=3, l_max=L)
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        pl = Planet.with_radius(R, Om, grid_type=kind, nlat=32, nlon=64, grid_resolution=3,
                                l_max=L, product_quadrature="fine")
    op = OPlanet(radius=R, omega=Om, basis="stable", dense=False, **kw)
    rng = np.random.default_rng(7)
    a = _rand_coeffs(L, rng)
    b = _rand_coeffs(L, rng)
    for dealias in (True, False):
        got = _np(pl.so.jacobian_pseudospectral(torch_mod.as_tensor(a, device="cuda"),
                                                torch_mod.as_tensor(b, device="cuda"),
                                                dealias=dealias, return_spectral=True))
        ref = op.so.jacobian(a, b, op.psh, op.pcoslat, dealias=dealias)
        assert rel_l2(got, ref) < 1e-12


def test_velocity_and_gradient_match_oracle(torch_mod):
    from aeolus_amd.planet import Planet
    from oracle.models import Planet as OPlanet
    L, R, Om = 21, 6.371e6, 7.292e-5
    pl = Planet.with_radius(R, Om, grid_type="latlon", nlat=32, nlon=64, l_max=L)
    op = OPlanet(radius=R, omega=Om, grid_type="latlon", nlat=32, nlon=64, l_max=L)
    psi = _rand_coeffs(L, np.random.default_rng(3))
    u, v = pl.so.velocity_from_streamfunction(torch_mod.as_tensor(psi, device="cuda"))
    ur, vr = op.so.velocity_from_streamfunction(psi, op.sh, op.grid.point_latitudes)
    assert rel_l2(_np(u), ur) < 1e-13 and rel_l2(_np(v), vr) < 1e-13
    dx, dy = pl.so.grad_from_scalar(torch_mod.as_tensor(psi, device="cuda"))
    dxr, dyr = op.so.grad_from_scalar(psi, op.sh, np.cos(op.grid.point_latitudes))
    assert rel_l2(_np(dx), dxr) < 1e-13 and rel_l2(_np(dy), dyr) < 1e-13
