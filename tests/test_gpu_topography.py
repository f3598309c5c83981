"""``Topography.mountain`` on the device against the oracle's restatement of
``phy/topography.py:207-290`` (analytic Gaussian at the state points,
analysed once with the backend transform, optional ``l_cut``, 0.2 relative
projection-residual gate).

Bars: coefficients within 1e-13 relative L2 of the oracle (1e-12 on the
geodesic grid), the projection residual within 1e-10 relative (plus 1e-13
absolute, for residuals at round-off); an unrepresentable mountain is
rejected by both.
"""
import math
import os
import warnings

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

R = 6.371e6
OMEGA = 2.0 * math.pi / 86400.0
GRAVITY = 9.80616
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _planets(kind, L, nlat=0, nlon=0, res=0):
    from aeolus_amd.planet import Planet
    from oracle.models import Planet as OPlanet
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        pl = Planet.with_radius(R, OMEGA, grid_type=kind, nlat=nlat, nlon=nlon,
                                grid_resolution=res, l_max=L, product_quadrature="fine")
    op = OPlanet(l_max=L, radius=R, omega=OMEGA, grid_type=kind, nlat=nlat, nlon=nlon,
                 resolution=res, product_quadrature="coarse")
    return pl, op


def _tri(a):
    li, mi = np.tril_indices(a.shape[-1])
    return a[..., li, mi]


def _c3_mountain():
    g = np.load(os.path.join(GOLDEN, "c3.npz"))
    h, lat0, lon0, width = (float(v) for v in g["mountain"])
    return dict(height_m=h, lat_deg=lat0, lon_deg=lon0, width_deg=width)


@pytest.mark.parametrize("l_cut", [None, 56])
def test_mountain_c3_matches_oracle(l_cut):
    """Gauss 128 x 256, L = 85, 2 km mountain of 1000 km e-folding width at
    the seeded centre of the shallow-water case; l_cut = 2L/3 as the
    primitive-equation core passes it."""
    from aeolus_amd.physics import Topography
    from oracle.models import mountain
    print(f"case topography c3 L=85 l_cut={l_cut}", flush=True)
    kw = _c3_mountain()
    pl, op = _planets("latlon", 85, 128, 256)
    topo = Topography.mountain(pl, l_cut=l_cut, **kw)
    ref, rel_ref = mountain(op, l_cut=l_cut, return_residual=True, **kw)
    got = topo.elevation_lm.cpu().numpy()
    err = rel_l2(_tri(got), _tri(ref))
    print(f"  elevation rel-L2 {err:.3e}; residual {topo.parameters['projection_residual']:.6e}"
          f" vs {rel_ref:.6e}", flush=True)
    assert err < 1e-13
    if l_cut is not None:
        assert np.all(got[l_cut + 1:, :] == 0) and np.all(got[:, l_cut + 1:] == 0)
    assert np.all(np.triu(got, 1) == 0)          # padding m > l written as 0
    phi = topo.surface_geopotential_lm(GRAVITY).cpu().numpy()
    assert rel_l2(_tri(phi), _tri(GRAVITY * ref)) < 1e-13
    # without l_cut the 1000 km mountain is resolved at L = 85 under exact Gauss
    # quadrature and both residuals are round-off (~2e-14): compare absolutely
    res = topo.parameters["projection_residual"]
    assert abs(res - rel_ref) <= 1e-10 * rel_ref + 1e-13
    assert topo.parameters.get("l_cut") == l_cut


def test_mountain_geodesic_matches_oracle():
    """Geodesic res 4 (Voronoi weights), L = 21: the point-set transform path."""
    from aeolus_amd.physics import Topography
    from oracle.models import mountain
    print("case topography geodesic res 4 L=21", flush=True)
    kw = dict(height_m=1500.0, lat_deg=35.0, lon_deg=-60.0, width_deg=20.0)
    pl, op = _planets("geodesic", 21, res=4)
    topo = Topography.mountain(pl, **kw)
    ref, rel_ref = mountain(op, return_residual=True, **kw)
    err = rel_l2(_tri(topo.elevation_lm.cpu().numpy()), _tri(ref))
    print(f"  elevation rel-L2 {err:.3e}", flush=True)
    assert err < 1e-12
    assert abs(topo.parameters["projection_residual"] - rel_ref) <= 1e-9 * rel_ref


def test_mountain_rejections_agree_with_oracle():
    """A mountain too narrow for L = 21 is rejected by both (residual > 0.2)."""
    from aeolus_amd.physics import Topography, TopographyError
    from oracle.models import mountain
    print("case topography rejection L=21", flush=True)
    kw = dict(height_m=2000.0, lat_deg=10.0, lon_deg=20.0, width_deg=2.0)
    pl, op = _planets("latlon", 21, 32, 64)
    with pytest.raises(ValueError):
        mountain(op, **kw)
    with pytest.raises(TopographyError, match="projection residual"):
        Topography.mountain(pl, **kw)
    with pytest.raises(TopographyError, match="l_cut"):
        Topography.mountain(pl, l_cut=22, **dict(kw, width_deg=30.0))
