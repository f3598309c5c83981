"""Device diagnostics recorders against the oracle's rows: one row of
``DiagnosticsRecorder.record`` (run/bve/diagnostics.py:162-211) and of
``SWEDiagnosticsRecorder.record`` (run/swe/diagnostics.py:113-173), the CSV /
spectra files they write, and the SWE characteristic speed
(phy/shallow_water.py:470-484).

Bars: every number within 1e-11 relative of the oracle (grid extrema, the
compensated quadratures and the spectral sums); the round-trip residual
below 1e-12 where the quadrature is exact (Gauss) and within 1e-10 relative
of the oracle's on geodesic Voronoi weights, where it is a real residual.
"""
import csv
import math
import warnings

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

R = 6.371e6
OMEGA = 2.0 * math.pi / 86400.0


def _planets(kind, L, nlat=0, nlon=0, res=0):
    from aeolus_amd.planet import Planet
    from oracle.models import Planet as OPlanet
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        pl = Planet.with_radius(R, OMEGA, grid_type=kind, nlat=nlat, nlon=nlon,
                                grid_resolution=res, l_max=L, product_quadrature="fine")
    op = OPlanet(l_max=L, radius=R, omega=OMEGA, grid_type=kind, nlat=nlat, nlon=nlon,
                 resolution=res)
    return pl, op


def _close(a, b, tol=1e-11):
    if math.isnan(a) or math.isnan(b):
        return math.isnan(a) and math.isnan(b)
    return abs(a - b) <= tol * max(abs(a), abs(b), 1e-300)


@pytest.mark.parametrize("kind,L,nlat,nlon,res", [("latlon", 42, 64, 128, 0),
                                                  ("geodesic", 10, 0, 0, 3)])
def test_bve_recorder_matches_oracle(kind, L, nlat, nlon, res, tmp_path):
    import torch
    from aeolus_amd.run.diagnostics import CSV_COLUMNS, DiagnosticsRecorder
    from aeolus_amd.run.initial_conditions import spectrum_l15
    from oracle.diagnostics import bve_record_row
    print(f"case bve recorder {kind} L={L}", flush=True)
    pl, op = _planets(kind, L, nlat, nlon, res)
    ls = pl.grid.cfl_length_scale
    rec = DiagnosticsRecorder(pl.sh, pl.so, pl.grid, R, OMEGA, tmp_path, spectra_every=2,
                              roundtrip_every=1)
    dt = 600.0
    for step in range(3):
        z = spectrum_l15(L, np.random.default_rng(40 + step), rms=5e-5)
        row = rec.record(step * dt, torch.as_tensor(z, device="cuda"), dt, step)
        ref = bve_record_row(op, z, dt, ls)
        for k, v in ref.items():
            if k == "roundtrip_residual" and v < 1e-12:   # exact quadrature (Gauss)
                assert row[k] < 1e-12
            elif k == "roundtrip_residual":                 # Voronoi weights: a real residual
                assert _close(row[k], v, 1e-10), (row[k], v)
            else:
                assert _close(row[k], v), (k, row[k], v)
    rec.close()
    with open(tmp_path / "diagnostics" / "timeseries.csv") as f:
        rows = list(csv.reader(f))
    assert rows[0] == CSV_COLUMNS and len(rows) == 4
    sp = np.load(tmp_path / "diagnostics" / "spectra.npz")
    assert sp["energy_l"].shape == (2, L + 1) and sp["enstrophy_l"].shape == (2, L + 1)
    assert np.allclose(sp["times"], [0.0, 2 * dt])


@pytest.mark.parametrize("kind,L,nlat,nlon,res", [("latlon", 21, 32, 64, 0),
                                                  ("geodesic", 10, 0, 0, 3)])
def test_swe_recorder_and_char_speed_match_oracle(kind, L, nlat, nlon, res, tmp_path):
    import torch
    from aeolus_amd.physics import ShallowWaterModel, ShallowWaterState, Topography
    from aeolus_amd.run.initial_conditions import swe_spectra
    from aeolus_amd.run.swe_diagnostics import SWE_CSV_COLUMNS, SWEDiagnosticsRecorder
    from oracle.diagnostics import swe_record_row
    from oracle.models import SWEParams, mountain
    print(f"case swe recorder {kind} L={L}", flush=True)
    pl, op = _planets(kind, L, nlat, nlon, res)
    elev = mountain(op, height_m=2000.0, lat_deg=30.0, lon_deg=60.0,
                    width_deg=math.degrees(2e6 / R))
    topo = Topography(torch.as_tensor(elev, device="cuda"), preset="mountain")
    model = ShallowWaterModel(pl, gravity=9.80616, mean_depth=1e4, topography=topo)
    p = SWEParams(op, gravity=9.80616, mean_depth=1e4, elevation_lm=elev)
    ls = pl.grid.cfl_length_scale
    rec = SWEDiagnosticsRecorder(model, tmp_path)
    dt = 300.0
    for step in range(2):
        y = swe_spectra(L, R, 9.80616, seed=7 + step)
        st = ShallowWaterState(torch.as_tensor(y, device="cuda"))
        row = rec.record(step * dt, st, dt, step)
        ref = swe_record_row(p, y, dt, ls, terrain_max_m=model.surface_elevation_extrema[1])
        for k, v in ref.items():
            assert _close(row[k], v), (k, row[k], v)
        assert _close(model.max_characteristic_speed(st), ref["max_char_speed_ms"])
        lo, hi = model.characteristic_fields(st)["phi_total_extrema"]
        assert _close(lo, ref["phi_total_min"]) and _close(hi, ref["phi_total_max"])
    rec.close()
    with open(tmp_path / "diagnostics" / "timeseries.csv") as f:
        rows = list(csv.reader(f))
    assert rows[0] == SWE_CSV_COLUMNS and len(rows) == 3
