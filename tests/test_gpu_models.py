"""HIP BVE / SWE tendencies and RK4 trajectories against the CPU oracle.

Bars: tendencies within 1e-12 relative L2 of the oracle (up to L = 230:
the single-member kernel's second analysis slot and latitude passes at
L = 200, the generic stage kernel past the single-member kernel's range at
L = 230); the C1 BVE
trajectory (Gauss 12x24 / 13x25, L = 8, 144 RK4 steps of 600 s) within
1e-12 relative L2 (the parity bar of the plain FP64 path).
"""
import math
import warnings

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

R_C1 = 6.371e6
OMEGA = 2.0 * math.pi / 86400.0


def _np(t):
    return t.detach().cpu().numpy()


def _planets(kind, L, nlat=0, nlon=0, res=0, R=R_C1, Om=OMEGA):
    from aeolus_amd.planet import Planet
    from oracle.models import Planet as OPlanet
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        pl = Planet.with_radius(R, Om, grid_type=kind, nlat=nlat, nlon=nlon,
                                grid_resolution=res, l_max=L, product_quadrature="fine")
    op = OPlanet(l_max=L, radius=R, omega=Om, grid_type=kind, nlat=nlat, nlon=nlon,
                 resolution=res)
    return pl, op


@pytest.mark.parametrize("kind,L,nlat,nlon,res,nu", [
    ("latlon", 8, 12, 24, 0, 0.0), ("latlon", 21, 32, 64, 0, 1e5),
    ("latlon", 85, 128, 256, 0, 0.0), ("geodesic", 10, 0, 0, 3, 0.0),
    ("latlon", 200, 320, 640, 0, 1e5), ("latlon", 230, 352, 704, 0, 0.0),
    ("latlon", 42, 65, 128, 0, 0.0), ("latlon", 85, 129, 200, 0, 0.0),
    ("geodesic", 42, 0, 0, 5, 1e5)])
def test_bve_tendency_matches_oracle(kind, L, nlat, nlon, res, nu):
    import torch
    from aeolus_amd.physics import BarotropicState, BarotropicVorticity
    from aeolus_amd.run.initial_conditions import spectrum_l15
    from oracle.models import bve_tendency
    print(f"case bve tendency {kind} L={L}", flush=True)
    pl, op = _planets(kind, L, nlat, nlon, res)
    model = BarotropicVorticity(pl, viscosity=nu)
    z = spectrum_l15(L, np.random.default_rng(1), rms=5e-5)
    z[0, 0] = 3e-6  # exercise the pinned l = 0 row
    got = _np(model.tendency(BarotropicState(torch.as_tensor(z, device="cuda"))))
    ref = bve_tendency(op, z, nu=nu)
    assert rel_l2(got, ref) < 1e-12
    assert np.all(got[0, :] == 0)


def test_bve_c1_trajectory_matches_oracle():
    """C1: Gauss 12x24 / 13x25, L = 8, two Gaussian vortices, 144 x 600 s."""
    import torch
    from aeolus_amd.physics import BarotropicState, BarotropicVorticity
    from aeolus_amd.run import rk4_step
    from aeolus_amd.run.initial_conditions import gaussian_vortex_pair
    from oracle.models import bve_tendency, rk4_step as o_rk4
    pl, op = _planets("latlon", 8, 12, 24)
    model = BarotropicVorticity(pl)
    zg = gaussian_vortex_pair(op.grid.point_latitudes, op.grid.point_longitudes, R_C1, seed=1)
    z0 = op.sh.transform(zg)
    y_ref = z0.copy()
    for _ in range(144):
        y_ref = o_rk4(lambda y: bve_tendency(op, y), y_ref, 600.0)
    state = BarotropicState(torch.as_tensor(z0, device="cuda"))
    for _ in range(144):
        state = rk4_step(model, state, 0.0, 600.0)
    assert rel_l2(_np(state.coeffs), y_ref) < 1e-12
    fused = model.rk4(torch.as_tensor(z0, device="cuda"), 600.0, 144)
    assert rel_l2(_np(fused), y_ref) < 1e-12


@pytest.mark.parametrize("L,nlat,nlon", [(42, 65, 128), (85, 129, 200), (200, 320, 640)])
def test_bve_fused_rk4_matches_oracle_short(L, nlat, nlon):
    """The fused RK4 step (single-member kernel, split heavy columns, the
    per-row grid kernel on power-of-two rows; an equator row on odd grids)
    against the oracle's RK4, 6 steps of 300 s, within 1e-12."""
    import torch
    from aeolus_amd.physics import BarotropicVorticity
    from aeolus_amd.run.initial_conditions import spectrum_l15
    from oracle.models import bve_tendency, rk4_step as o_rk4
    print(f"case bve fused rk4 L={L} {nlat}x{nlon}", flush=True)
    pl, op = _planets("latlon", L, nlat, nlon)
    model = BarotropicVorticity(pl, viscosity=1e5)
    z0 = spectrum_l15(L, np.random.default_rng(7), rms=5e-5)
    y_ref = z0.copy()
    for _ in range(6):
        y_ref = o_rk4(lambda y: bve_tendency(op, y, nu=1e5), y_ref, 300.0)
    got = model.rk4(torch.as_tensor(z0, device="cuda"), 300.0, 6)
    assert rel_l2(_np(got), y_ref) < 1e-12


@pytest.mark.parametrize("kind,L,nlat,nlon,res,nu4,topo", [
    ("latlon", 15, 32, 64, 0, 0.0, False), ("latlon", 21, 32, 64, 0, 1e14, True),
    ("latlon", 85, 128, 256, 0, 0.0, True), ("geodesic", 10, 0, 0, 3, 0.0, True),
    ("latlon", 170, 256, 512, 0, 0.0, False), ("geodesic", 42, 0, 0, 5, 0.0, True)])
def test_swe_tendency_matches_oracle(kind, L, nlat, nlon, res, nu4, topo):
    import torch
    from aeolus_amd.physics import ShallowWaterModel, Topography
    from aeolus_amd.run.initial_conditions import swe_spectra
    from oracle.models import SWEParams, mountain, swe_tendency
    print(f"case swe tendency {kind} L={L}", flush=True)
    pl, op = _planets(kind, L, nlat, nlon, res)
    topog = None
    elev = None
    if topo:
        elev = mountain(op, height_m=2000.0, lat_deg=30.0, lon_deg=60.0,
                        width_deg=math.degrees(2e6 / R_C1))
        topog = Topography(torch.as_tensor(elev, device="cuda"), preset="mountain")
    model = ShallowWaterModel(pl, gravity=9.80616, mean_depth=1e4, hyperdiffusion_nu4=nu4,
                              topography=topog)
    p = SWEParams(op, gravity=9.80616, mean_depth=1e4, nu4=nu4, elevation_lm=elev)
    y = swe_spectra(L, R_C1, 9.80616, seed=3)
    got = _np(model.tendency(torch.as_tensor(y, device="cuda")))
    ref = swe_tendency(p, y)
    for f in range(3):
        assert rel_l2(got[f], ref[f]) < 1e-12, f
    assert np.all(got[:, 0, :] == 0)


def test_swe_rk4_matches_oracle_short():
    import torch
    from aeolus_amd.physics import ShallowWaterModel
    from aeolus_amd.run import rk4_step_array
    from aeolus_amd.run.initial_conditions import swe_spectra
    from oracle.models import SWEParams, rk4_step_array as o_rk4, swe_tendency
    L = 21
    pl, op = _planets("latlon", L, 32, 64)
    model = ShallowWaterModel(pl, mean_depth=1e4)
    p = SWEParams(op, mean_depth=1e4)
    y0 = swe_spectra(L, R_C1, 9.80616, seed=4)
    y_ref = y0.copy()
    for _ in range(10):
        y_ref = o_rk4(lambda y: swe_tendency(p, y), y_ref, 300.0)
    y = torch.as_tensor(y0, device="cuda")
    for _ in range(10):
        y = rk4_step_array(model.tendency, y, 0.0, 300.0)
    for f in range(3):
        assert rel_l2(_np(y)[f], y_ref[f]) < 1e-12
    # composed path (stage validator) agrees with the fused path
    y2 = torch.as_tensor(y0, device="cuda")
    for _ in range(10):
        y2 = rk4_step_array(model.tendency, y2, 0.0, 300.0,
                            stage_validator=lambda s: None)
    assert rel_l2(_np(y2), _np(y)) < 1e-13


@pytest.mark.parametrize("L,nlat,nlon,nmem", [(42, 64, 128, 1), (42, 64, 128, 3),
                                               (42, 64, 128, 4), (42, 64, 128, 9),
                                               (85, 128, 256, 2), (170, 256, 512, 3),
                                               (200, 320, 640, 2), (255, 384, 768, 2)])
def test_bve_batch_members_match_single_runs(L, nlat, nlon, nmem):
    """Every member of a batched step (the generic stage kernel) equals the
    single-member step (the dedicated single-member kernel, whose heavy
    columns run as two degree halves whose Legendre sums the grid stage adds)
    to rounding: the per-degree analysis sums are the same, the synthesis sums
    of a split column are associated per degree half.  L = 200 takes the
    single-member kernel's second analysis slot (a half with more than 64
    analysed degrees) and more than one synthesis latitude pass (nh > 128);
    L = 255 has columns with more analysed degrees than the single-member
    kernel's two slots hold."""
    import torch
    from aeolus_amd.physics import BarotropicVorticity
    from aeolus_amd.physics.barotropic import BarotropicBatch
    from aeolus_amd.run.initial_conditions import spectrum_l15
    print(f"case bve batch L={L} nmem={nmem}", flush=True)
    pl, _ = _planets("latlon", L, nlat, nlon)
    single = BarotropicVorticity(pl)
    batch = BarotropicBatch(pl, nmem)
    zs = np.stack([spectrum_l15(L, np.random.default_rng(100 + k), rms=5e-5)
                   for k in range(nmem)])
    out = batch.rk4(torch.as_tensor(zs, device="cuda"), 600.0, 5)
    tend = batch.tendency(BarotropicState_(torch.as_tensor(zs, device="cuda")))
    for k in range(nmem):
        ref = single.rk4(torch.as_tensor(zs[k], device="cuda"), 600.0, 5)
        assert rel_l2(_np(out[k]), _np(ref)) < 1e-13, k
        tk = single.tendency(BarotropicState_(torch.as_tensor(zs[k], device="cuda")))
        assert rel_l2(_np(tend[k]), _np(tk)) < 1e-13, k


def BarotropicState_(c):
    from aeolus_amd.physics import BarotropicState
    return BarotropicState(c)


def test_swe_validate_state_rejections():
    """SWE hard validation (``phy/shallow_water.py:428-454``): NaN/Inf, each
    prognostic's monopole, non-positive thickness on the state or product
    sampling; a valid state passes."""
    import torch
    from aeolus_amd.physics import ShallowWaterModel, ShallowWaterState, ShallowWaterStateError
    from aeolus_amd.run.initial_conditions import swe_spectra
    pl, _ = _planets("latlon", 21, 32, 64)
    m = ShallowWaterModel(pl, gravity=9.80616, mean_depth=1e4)
    y0 = torch.as_tensor(swe_spectra(21, R_C1, 9.80616, seed=4), device="cuda")
    m.validate_state(ShallowWaterState(y0))
    cases = []
    y = y0.clone(); y[0, 5, 3] = float("nan"); cases.append((y, "NaN/Inf"))
    for f, name in enumerate(("zeta", "delta", "phi")):
        y = y0.clone(); y[f, 0, 0] = 1.0; cases.append((y, f"{name} monopole"))
    y = y0.clone(); y[2, 1, 0] = -3.0 * m.phi0; cases.append((y, "not strictly positive"))
    for y, msg in cases:
        with pytest.raises(ShallowWaterStateError, match=msg):
            m.validate_state(ShallowWaterState(y))
    lo, hi = m.total_geopotential_extrema(ShallowWaterState(y0))
    assert 0.0 < lo < m.phi0 < hi
