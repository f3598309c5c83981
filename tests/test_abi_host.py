"""C ABI and host-side logic on the CPU (no kernel launches).

* the library loads and exports every function include/aeolus_hip.h declares;
* argument validation returns the documented status before touching the
  device (the reference's "no silent fallback" ValueError contract);
* backend seam contract (tests/test_spherical_backend.py:44-120) with a
  CPU-resident geometry;
* host geometry equals the oracle's; seeded initial conditions hit their
  stated norms.
"""
import ctypes
import math
import re
from pathlib import Path

import numpy as np
import pytest
import torch

from aeolus_amd import _native as nat

HEADER = Path(__file__).resolve().parents[1] / "include" / "aeolus_hip.h"


def _declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(aeo_\w+)\s*\(", text, re.M)))


def test_every_declared_symbol_is_exported():
    lib = nat.lib()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) <= set(nat._SIGNATURES)
    assert lib.aeo_version() == 1


def test_invalid_arguments_are_rejected_on_the_host():
    lib = nat.lib()
    out = ctypes.c_void_p()
    x = np.array([0.5, -0.5])
    w = np.array([1.0, 1.0])
    c = np.array([0.8, 0.8])
    rc = lib.aeo_plan_create_gauss(1, 8, 2, nat.host_ptr(x), nat.host_ptr(w), nat.host_ptr(c), 1,
                                   ctypes.byref(out))
    assert rc == nat.AEO_ERR_INVALID and b"nlat" in lib.aeo_last_error()
    rc = lib.aeo_plan_create_gauss(2, 8, 2, None, nat.host_ptr(w), nat.host_ptr(c), 1,
                                   ctypes.byref(out))
    assert rc == nat.AEO_ERR_INVALID
    bad = np.array([0.5, -0.4])   # not symmetric
    rc = lib.aeo_plan_create_gauss(2, 8, 2, nat.host_ptr(bad), nat.host_ptr(w), nat.host_ptr(c), 1,
                                   ctypes.byref(out))
    assert rc == nat.AEO_ERR_INVALID and b"symmetric" in lib.aeo_last_error()
    rc = lib.aeo_plan_create_gauss(2, 4096, 2, nat.host_ptr(x), nat.host_ptr(w), nat.host_ptr(c), 1,
                                   ctypes.byref(out))
    assert rc == nat.AEO_ERR_UNSUPPORTED
    rc = lib.aeo_spec_op(99, 4, 1.0, 16, 16, 1, None)
    assert rc == nat.AEO_ERR_UNSUPPORTED
    with pytest.raises(ValueError, match="no silent fallback"):
        nat.check(rc, "spec_op")
    assert lib.aeo_spec_op(0, 4, 1.0, None, 16, 1, None) == nat.AEO_ERR_INVALID
    assert lib.aeo_spec_op(0, 4, -1.0, 16, 16, 1, None) == nat.AEO_ERR_INVALID
    assert lib.aeo_analysis(None, 16, 1, 16, None) == nat.AEO_ERR_INVALID
    assert lib.aeo_model_rk4(None, 16, 1.0, 1, None, 16, None) == nat.AEO_ERR_INVALID
    assert lib.aeo_bve_create(None, 1.0, 0.0, 0.0, ctypes.byref(out)) == nat.AEO_ERR_INVALID
    assert lib.aeo_swe_create(None, 1.0, 0.0, -1.0, 1.0, 0.0, None,
                              ctypes.byref(out)) == nat.AEO_ERR_INVALID
    assert lib.aeo_model_profile(None, 0, 1) == nat.AEO_ERR_INVALID
    assert lib.aeo_model_enable_adjoint(None) == nat.AEO_ERR_INVALID
    assert lib.aeo_model_rk4_adjoint(None, 16, 1.0, 1, 16, 16, None) == nat.AEO_ERR_INVALID


class _FakeSH:
    l_max = 8
    device = torch.device("cpu")


def test_backend_contract_on_cpu_geometry():
    from aeolus_amd.numerics.latlon_grid import GaussLatLonGridGeometry
    from aeolus_amd.numerics.spherical_backend import LatLonBackend, PointSetBackend
    g = GaussLatLonGridGeometry(12, 24, 6.371e6, device=torch.device("cpu"))
    be = LatLonBackend(g, _FakeSH())
    assert be.supported_product_quadratures() == ("coarse", "fine")
    with pytest.raises(ValueError, match="no silent fallback"):
        be.product_space("bogus")
    ps = be.product_space("coarse")
    assert ps is be.product_space("coarse")        # cached by identity
    assert ps.sh is be.sh and ps.geometry is None
    assert ps.label == "latlon-gauss-12x24-state"
    assert float(ps.coslat.min()) >= 1e-8
    with pytest.raises(ValueError):
        LatLonBackend(object(), _FakeSH())
    pb = PointSetBackend(g, _FakeSH())
    assert pb.supported_product_quadratures() == ("coarse",)
    with pytest.raises(ValueError, match="no silent fallback"):
        pb.product_space("fine")
    assert abs(float(g.solid_angle_weights.sum()) - 4 * math.pi) < 1e-13


def test_host_geodesic_geometry_matches_oracle():
    from aeolus_amd.numerics.geodesic_grid import GeodesicGridGeometry
    from oracle.grids import GeodesicGrid
    for res in (0, 2, 4):
        g = GeodesicGridGeometry(res, 6.371e6, device=torch.device("cpu"))
        o = GeodesicGrid(res, 6.371e6)
        assert g.n_points == o.n_points
        assert np.abs(g.points - o.points).max() < 1e-8          # metres
        assert np.array_equal(g.faces, o.faces)
        assert np.abs(g.cell_areas_np - o.cell_areas).max() < 1e-9 * o.cell_areas.max()


def test_planetary_radius():
    from aeolus_amd.planet import PlanetaryParameters
    p = PlanetaryParameters.from_earth_like(day_hours=24.0)
    assert abs(p.radius - 6367354.69) < 0.01
    assert p.angular_velocity == 2 * math.pi / 86400.0


def test_initial_conditions_norms():
    from aeolus_amd.run.initial_conditions import field_rms, power_per_degree, spectrum_l15, swe_spectra
    a = spectrum_l15(21, np.random.default_rng(0), rms=5e-5)
    assert abs(field_rms(a) - 5e-5) < 1e-18
    assert np.all(a[0] == 0) and np.all(a[:, 0].imag == 0)
    li, mi = np.indices(a.shape)
    assert np.all(a[mi > li] == 0)
    R = 6.371e6
    y = swe_spectra(85, R, 9.80616, seed=3)
    l = np.arange(86.0)
    w = np.zeros(86)
    w[1:] = R ** 2 / (l[1:] * (l[1:] + 1))
    u2 = np.sum(w * (power_per_degree(y[0]) + power_per_degree(y[1]))) / (4 * math.pi)
    assert abs(math.sqrt(u2) - 20.0) < 1e-12
    assert abs(field_rms(y[1]) / field_rms(y[0]) - 0.1) < 1e-14
    assert abs(field_rms(y[2]) - 9.80616 * 100.0) < 1e-10
