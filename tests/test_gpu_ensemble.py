"""Uneven-L ensemble: graph-captured step == eager step (bitwise) == per-member
runs (bitwise for split precision; to rounding in FP64, where a single member
runs the dedicated single-member kernel), the autotuner picks a tile that
runs, odd and even step counts."""
import math
import warnings

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

R = 6.371e6
OM = 2.0 * math.pi / 86400.0
CLASSES = [(21, 32, 64, 5), (42, 64, 128, 3), (85, 128, 256, 2)]


def _planet(L, nlat, nlon):
    from aeolus_amd.planet import Planet
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return Planet.with_radius(R, OM, grid_type="latlon", nlat=nlat, nlon=nlon, l_max=L)


@pytest.mark.parametrize("precision", [0, 6])
def test_ensemble_graph_matches_eager_and_single(precision, tmp_path, monkeypatch):
    import torch
    from aeolus_amd.physics import BarotropicVorticity
    from aeolus_amd.run.ensemble import Ensemble
    from aeolus_amd.run.initial_conditions import spectrum_l15
    monkeypatch.setenv("AEOLUS_AMD_CACHE", str(tmp_path))
    print(f"case ensemble precision={precision}", flush=True)
    planets = [_planet(L, a, b) for L, a, b, _ in CLASSES]
    ens = Ensemble([(pl, n) for pl, (_, _, _, n) in zip(planets, CLASSES)], precision=precision)
    states = [torch.as_tensor(np.stack([spectrum_l15(L, np.random.default_rng(10 * L + k),
                                                     rms=5e-5) for k in range(n)]), device="cuda")
              for L, _, _, n in CLASSES]
    tuning = ens.autotune(states, 300.0, steps=2)
    assert all(t["mb"] in (1, 2, 4, 8) or (t["mb"] == 0 and precision > 0) for t in tuning)
    if precision:
        # the GEMM form was among the candidates of every split class
        assert all(any(k.startswith("gemm") for k in t["ms_per_step"]) for t in tuning)
    for nsteps in (4, 5):
        g = ens.rk4(states, 300.0, nsteps, use_graph=True)
        e = ens.rk4(states, 300.0, nsteps, use_graph=False)
        for a, b in zip(g, e):
            assert torch.equal(a, b)
    for (L, _, _, n), pl, y0, y in zip(CLASSES, planets, states, g):
        single = BarotropicVorticity(pl)
        single.set_precision(precision)
        for k in range(n):
            ref = single.rk4(y0[k], 300.0, 5)
            if precision:
                assert torch.equal(y[k], ref), (L, k)
            else:
                # FP64 single members run the dedicated single-member kernel
                # (split columns' synthesis sums associated per degree half)
                err = float((y[k] - ref).abs().pow(2).sum().sqrt() / ref.abs().pow(2).sum().sqrt())
                assert err < 1e-13, (L, k, err)
    # the cached choice is reused
    again = ens.autotune(states, 300.0, steps=2)
    assert all(t.get("cached") for t in again)


def test_graph_recaptured_after_shared_plan_tables_rebuilt():
    """Another model on the same planet switches the shared plan's digit tables
    6 -> 7 -> 6 slices (rebuilt at new addresses); the ensemble's next graph
    replay must use the rebuilt tables: graph == eager, bitwise."""
    import torch
    from aeolus_amd.physics import BarotropicVorticity
    from aeolus_amd.run.ensemble import Ensemble
    from aeolus_amd.run.initial_conditions import spectrum_l15
    print("case ensemble graph after plan table rebuild", flush=True)
    pl = _planet(42, 64, 128)
    ens = Ensemble([(pl, 3)], precision=6)
    y0 = torch.as_tensor(np.stack([spectrum_l15(42, np.random.default_rng(500 + k), rms=5e-5)
                                   for k in range(3)]), device="cuda")
    first = ens.rk4([y0], 300.0, 4, use_graph=True)[0]
    other = BarotropicVorticity(pl)
    other.set_precision(7)
    other.set_precision(6)
    g = ens.rk4([y0], 300.0, 4, use_graph=True)[0]
    e = ens.rk4([y0], 300.0, 4, use_graph=False)[0]
    assert torch.equal(g, e)
    assert torch.equal(g, first)
