"""Pins of the oracle's grids and transforms (CPU).

Known answers from the reference's own tests: Gauss weights sum to 4 pi
(tests/test_latlon_transform.py:51), 3/2-rule product sizes
(tests/test_latlon_backend.py:90-101), Gauss round trip / single-mode delta /
a00 = sqrt(4 pi) (tests/test_latlon_transform.py:99-121), geodesic round
trip envelope 5e-2 at res 4 / L = 21 (tests/test_backend_parity.py:81).
Independent checks: geodesic point count 10*4^r + 2, areas summing to
4 pi R^2, SciPy rotation, FFT+Legendre == dense reference-formula transform.
"""
import math

import numpy as np
import pytest

from oracle.grids import GaussGrid, GeodesicGrid, product_grid_dims
from oracle.transforms import DenseTransform, GaussTransform


def _rand(L, rng):
    a = np.zeros((L + 1, L + 1), dtype=np.complex128)
    for l in range(L + 1):
        a[l, 0] = rng.standard_normal()
        a[l, 1:l + 1] = rng.standard_normal(l) + 1j * rng.standard_normal(l)
    return a


def test_gauss_weights_and_layout():
    g = GaussGrid(24, 48, radius=6.371e6)
    assert abs(g.weights.sum() - 4 * math.pi) / (4 * math.pi) < 1e-13
    assert np.all(np.diff(g.latitudes) < 0)
    assert g.longitudes[0] == 0.0 and g.longitudes[-1] < 2 * math.pi
    assert np.array_equal(g.x, -g.x[::-1])                  # leggauss symmetrises exactly
    assert np.array_equal(g.gl_weights, g.gl_weights[::-1])


@pytest.mark.parametrize("L,n,m,exp", [(8, 12, 24, (13, 25)), (85, 128, 256, (128, 256)),
                                       (170, 256, 512, (256, 512)), (21, 32, 64, (32, 64)),
                                       (15, 32, 64, (32, 64)), (10, 24, 48, (24, 48))])
def test_product_grid_three_halves_rule(L, n, m, exp):
    assert product_grid_dims(L, n, m) == exp


def test_geodesic_geometry():
    from scipy.spatial.transform import Rotation
    from oracle.grids import _ICO_V, _rotation_x
    assert np.abs(Rotation.from_rotvec([1e-2, 0, 0]).as_matrix() - _rotation_x(1e-2)).max() < 1e-16
    for r in range(5):
        g = GeodesicGrid(r, 6.371e6)
        assert g.n_points == 10 * 4 ** r + 2
        assert np.allclose(np.linalg.norm(g.unit_points, axis=1), 1.0, atol=1e-15)
    g = GeodesicGrid(4, 6.371e6)
    assert abs(g.cell_areas.sum() - 4 * math.pi * 6.371e6 ** 2) / (4 * math.pi * 6.371e6 ** 2) < 1e-10
    assert abs(g.voronoi_weights.sum() - 4 * math.pi) < 1e-12
    assert np.allclose(g.unit_points[:12], (_rotation_x(1e-2) @ _ICO_V.T).T, atol=1e-15)


@pytest.mark.parametrize("nlat,nlon,L", [(12, 24, 8), (13, 25, 8), (24, 48, 10), (32, 64, 21),
                                         (64, 128, 42)])
def test_fft_legendre_equals_dense_reference_transform(nlat, nlon, L):
    rng = np.random.default_rng(L)
    g = GaussGrid(nlat, nlon)
    gt = GaussTransform(g, L)
    dt = DenseTransform(g.point_latitudes, g.point_longitudes, L, g.weights, basis="reference")
    a = _rand(L, rng)
    f_d = dt.inv_transform(a)
    assert np.linalg.norm(gt.inv_transform(a) - f_d) / np.linalg.norm(f_d) < 1e-13
    f = rng.standard_normal(g.n_points)
    a_d = dt.transform(f)
    assert np.linalg.norm(gt.transform(f) - a_d) / np.linalg.norm(a_d) < 1e-13


def test_gauss_known_answers():
    L = 10
    g = GaussGrid(24, 48, 6.371e6)
    t = GaussTransform(g, L)
    a = _rand(L, np.random.default_rng(11)) * 1e-5
    a[:, 0] = a[:, 0].real
    back = t.transform(t.inv_transform(a))
    assert np.abs(back - a).max() / np.abs(a).max() < 1e-13
    d = np.zeros_like(a)
    d[3, 2] = 1.0 - 0.5j
    back = t.transform(t.inv_transform(d))
    assert abs(back[3, 2] - (1.0 - 0.5j)) < 1e-13
    back[3, 2] = 0
    assert np.abs(back).max() < 1e-13
    a00 = t.transform(np.ones(g.n_points))[0, 0]
    assert abs(a00 - math.sqrt(4 * math.pi)) / math.sqrt(4 * math.pi) < 1e-13


def test_synthesis_ignores_imag_l0_and_padding():
    L = 8
    g = GaussGrid(12, 24)
    t = GaussTransform(g, L)
    dt = DenseTransform(g.point_latitudes, g.point_longitudes, L, g.weights)
    a = _rand(L, np.random.default_rng(2))
    b = a.copy()
    b[:, 0] += 1j * 7.0
    b[0, 5] = 3.0   # padding (m > l)
    for tr in (t, dt):
        assert np.array_equal(tr.inv_transform(a), tr.inv_transform(b)) or \
            np.abs(tr.inv_transform(a) - tr.inv_transform(b)).max() < 1e-14


def test_geodesic_roundtrip_envelope():
    g = GeodesicGrid(4, 1.0)
    t = DenseTransform(g.latitudes, g.longitudes, 21, g.voronoi_weights)
    a = _rand(21, np.random.default_rng(4))
    a[:, 0] = a[:, 0].real
    back = t.transform(t.inv_transform(a))
    err = np.linalg.norm(back - a) / np.linalg.norm(a)
    assert 1e-4 < err < 5e-2   # Voronoi quadrature is approximate, not exact
