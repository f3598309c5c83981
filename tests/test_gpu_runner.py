"""Run drivers on the device: state-adaptive CFL integration with the
reference's snapshot contracts and capsule files (run/bve/runner.py:32-207,
run/swe/runner.py:29-135).

Bars: snapshot times exactly the requested schedule (count mode); the final
state within 1e-12 relative L2 of the oracle integrated with the same
accepted-step sequence (read back from the diagnostics CSV); one CSV row per
accepted step plus the initial row.
"""
import csv
import math
import warnings

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

R = 6.371e6
OMEGA = 2.0 * math.pi / 86400.0


def _planets(L, nlat, nlon):
    from aeolus_amd.planet import Planet
    from oracle.models import Planet as OPlanet
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        pl = Planet.with_radius(R, OMEGA, grid_type="latlon", nlat=nlat, nlon=nlon, l_max=L,
                                product_quadrature="fine")
    return pl, OPlanet(l_max=L, radius=R, omega=OMEGA, grid_type="latlon", nlat=nlat, nlon=nlon)


def _rows(path):
    with open(path) as f:
        return list(csv.DictReader(f))


def test_run_bve_count_mode_matches_oracle_steps(tmp_path):
    from aeolus_amd.run import count_snapshot_times, run_bve
    from aeolus_amd.run.initial_conditions import spectrum_l15
    from oracle.models import bve_tendency, rk4_step
    print("case run_bve count L=21", flush=True)
    L = 21
    pl, op = _planets(L, 32, 64)
    z0 = spectrum_l15(L, np.random.default_rng(3), rms=5e-5)
    t_end_days = 0.0531
    sched = count_snapshot_times(3, t_end_days * 86400.0)
    assert run_bve(pl, z0, None, t_end_days, tmp_path, 1e5, snapshot_times=sched) == 0
    times = np.load(tmp_path / "bve_snapshot_times.npy")
    assert times.tolist() == sched
    coeffs = np.load(tmp_path / "vorticity_coeffs.npy")
    assert coeffs.shape == (3, L + 1, L + 1) and coeffs.dtype == np.complex128
    assert np.load(tmp_path / "vorticity_grid.npy").shape == (3, 32 * 64)
    rows = _rows(tmp_path / "diagnostics" / "timeseries.csv")
    dts = [float(r["dt_s"]) for r in rows[1:]]
    assert len(rows) == 1 + len(dts) and float(rows[-1]["time_s"]) == sched[-1]
    # the CFL ceiling of each step: 0.5 * length scale / max speed of the previous row
    ls = pl.grid.cfl_length_scale
    for prev, dt in zip(rows[:-1], dts):
        assert dt <= 0.5 * ls / float(prev["max_speed_ms"]) * (1 + 1e-12)
    y = z0
    for dt in dts:
        y = rk4_step(lambda s: bve_tendency(op, s, nu=1e5), y, dt)
    assert rel_l2(coeffs[-1], y) < 1e-12
    assert (tmp_path / "diagnostics" / "spectra.npz").exists()


def test_run_swe_interval_mode_with_validation(tmp_path):
    import torch
    from aeolus_amd.physics import ShallowWaterModel, ShallowWaterState
    from aeolus_amd.run import run_swe
    from aeolus_amd.run.initial_conditions import swe_spectra
    from oracle.models import SWEParams, rk4_step_array, swe_tendency
    print("case run_swe interval L=15", flush=True)
    L = 15
    pl, op = _planets(L, 32, 64)
    model = ShallowWaterModel(pl, gravity=9.80616, mean_depth=1e4)
    y0 = swe_spectra(L, R, 9.80616, seed=5)
    st = ShallowWaterState(torch.as_tensor(y0, device="cuda"))
    t_end_days, dts = 0.02, 864.0
    assert run_swe(model, st, dts, t_end_days, tmp_path) == 0
    times = np.load(tmp_path / "swe_snapshot_times.npy")
    assert times[0] == 0.0 and np.allclose(times, np.arange(len(times)) * dts)
    snaps = np.load(tmp_path / "swe_coeffs.npy")
    assert snaps.shape == (len(times), 3, L + 1, L + 1)
    rows = _rows(tmp_path / "diagnostics" / "timeseries.csv")
    p = SWEParams(op, gravity=9.80616, mean_depth=1e4)
    y, t = y0, 0.0
    for r in rows[1:]:
        y = rk4_step_array(lambda s: swe_tendency(p, s), y, float(r["dt_s"]))
        t += float(r["dt_s"])
        if any(abs(t - s) <= 1e-6 * dts for s in times[1:]):
            k = int(np.argmin(np.abs(times - t)))
            assert rel_l2(snaps[k], y) < 1e-12, k
    assert all(float(r["h_min_m"]) > 0 for r in rows)


def test_run_swe_rejects_invalid_initial_state(tmp_path):
    import torch
    from aeolus_amd.physics import ShallowWaterModel, ShallowWaterState, ShallowWaterStateError
    from aeolus_amd.run import run_swe
    L = 10
    pl, _ = _planets(L, 16, 32)
    model = ShallowWaterModel(pl, gravity=9.80616, mean_depth=1e4)
    y = np.zeros((3, L + 1, L + 1), dtype=np.complex128)
    y[2, 0, 0] = 1.0     # nonzero geopotential monopole
    with pytest.raises(ShallowWaterStateError):
        run_swe(model, ShallowWaterState(torch.as_tensor(y, device="cuda")), 600.0, 0.01, tmp_path)
