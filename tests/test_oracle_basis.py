"""Pins of the oracle's basis constructions (CPU).

* closed forms Y00, Y10, Y11, Y20;
* 50-digit mpmath.spherharm spot values at L = 170 on the polar-most and a
  mid-latitude node of the 256-row Gauss grid;
* Gram matrix = I under exact Gauss quadrature;
* reference formula (cu/sh_matrix.cu) == stable recurrence to 2e-13 for L <= 85,
  and the reference formula's breakdown at L = 170 (the reason the oracle
  uses the stable recurrence there).
"""
import math

import numpy as np
import pytest

from oracle.basis import (flat_index, lm_indices, n_basis, normalized_legendre,
                          reference_basis, stable_basis)
from oracle.grids import GaussGrid


def test_flat_index_order():
    li, mi = lm_indices(3)
    assert list(zip(li, mi))[:6] == [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]
    assert n_basis(170) == 14706
    assert flat_index(170, 170) == n_basis(170) - 1


@pytest.mark.parametrize("builder", [reference_basis, stable_basis])
def test_closed_forms(builder):
    rng = np.random.default_rng(0)
    lat = rng.uniform(-1.5, 1.5, 40)
    lon = rng.uniform(0, 2 * np.pi, 40)
    Y = builder(lat, lon, 3)
    s, c = np.sin(lat), np.cos(lat)
    assert np.allclose(Y[:, 0], 1 / math.sqrt(4 * math.pi), atol=1e-15, rtol=0)
    assert np.allclose(Y[:, 1], math.sqrt(3 / (4 * math.pi)) * s, atol=1e-15, rtol=0)
    assert np.allclose(Y[:, 2], -math.sqrt(3 / (8 * math.pi)) * c * np.exp(1j * lon),
                       atol=1e-15, rtol=0)
    assert np.allclose(Y[:, 3], math.sqrt(5 / (16 * math.pi)) * (3 * s * s - 1), atol=1e-15,
                       rtol=0)


def test_mpmath_spot_values_at_L170():
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 50
    g = GaussGrid(256, 512)
    # convention guard: mpmath's Y_1^1 carries the Condon-Shortley phase
    th = mp.mpf("0.7")
    assert abs(complex(mp.spherharm(1, 1, th, 0)) + math.sqrt(3 / (8 * math.pi)) * math.sin(0.7)) < 1e-15
    P = normalized_legendre(g.x[[0, 70]], 170)
    for j, row in ((0, 0), (70, 1)):
        theta = mp.acos(mp.mpf(float(g.x[j])))
        for l, m in [(0, 0), (10, 3), (100, 37), (170, 0), (170, 1), (170, 85), (150, 150),
                     (170, 170)]:
            exact = mp.re(mp.spherharm(l, m, theta, 0))
            got = P[flat_index(l, m), row]
            # forward-recurrence rounding grows about linearly in l (~500 eps at
            # l = 170 next to the pole); a wrong term or index is an O(1) error
            assert abs(got - float(exact)) <= 5e-13 * max(1.0, abs(float(exact))), (j, l, m)


def test_stable_recurrence_is_finite_and_underflows_gracefully():
    g = GaussGrid(256, 512)
    P = normalized_legendre(g.x, 170)
    assert np.all(np.isfinite(P))
    # polar-most node: P_170,170 ~ sin^170 ~ 1e-345 underflows to (signed) zero
    assert abs(P[flat_index(170, 170), 0]) < 1e-300


def test_gram_is_identity_under_gauss_quadrature():
    L = 42
    g = GaussGrid(64, 128)
    Y = stable_basis(g.point_latitudes, g.point_longitudes, L)
    G = Y.conj().T @ (g.weights[:, None] * Y)
    assert np.abs(G - np.eye(G.shape[0])).max() < 1e-13


@pytest.mark.parametrize("L,tol", [(8, 1e-14), (21, 3e-14), (42, 1e-13), (85, 3e-13)])
def test_reference_formula_equals_stable_recurrence(L, tol):
    rng = np.random.default_rng(L)
    lat = np.concatenate([rng.uniform(-1.55, 1.55, 60), [1.5607963, -1.5607963, 0.0]])
    lon = rng.uniform(0, 2 * np.pi, lat.size)
    assert np.abs(reference_basis(lat, lon, L) - stable_basis(lat, lon, L)).max() < tol


def test_reference_formula_breaks_down_at_L170():
    g = GaussGrid(256, 512)
    lat = g.latitudes[40:128:8]   # mid and low latitudes: unnormalised P_mm overflows there
    with np.errstate(over="ignore", invalid="ignore"):
        Yr = reference_basis(lat, np.zeros_like(lat), 170)
    Ys = stable_basis(lat, np.zeros_like(lat), 170)
    bad = ~np.isfinite(Yr) | (np.abs(Yr - Ys) > 1e-6)
    assert bad.any()
