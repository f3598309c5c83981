"""Split-precision dense Y-basis contractions on geodesic point sets (INT8
radix-254 digit MFMA, exact INT32 accumulation, FP64 recombination) against
the FP64 point path and the oracle's golden trajectory.

Bars: 6 slices within 1e-11 relative L2 of the FP64 tendency (BVE and each
SWE field), 7 slices within 1e-13, 4 slices coarser than 1e-12 but within
1e-6; the C2 trajectory (geodesic res 4, L = 21, 96 x 900 s) with 6 slices
within 1e-11 of the FP64 oracle's golden fixture.
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
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _planet(L, res):
    from aeolus_amd.planet import Planet
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return Planet.with_radius(R, OMEGA, grid_type="geodesic", grid_resolution=res, l_max=L,
                                  product_quadrature="fine")


@pytest.mark.parametrize("L,res", [(10, 3), (21, 4)])
def test_split_points_bve_tendency_matches_fp64(L, res):
    import torch
    from aeolus_amd.physics import BarotropicState, BarotropicVorticity
    from aeolus_amd.run.initial_conditions import spectrum_l15
    print(f"case split points bve L={L} res={res}", flush=True)
    m = BarotropicVorticity(_planet(L, res))
    z = torch.as_tensor(spectrum_l15(L, np.random.default_rng(L), rms=5e-5), device="cuda")
    ref = m.tendency(BarotropicState(z)).cpu().numpy()
    errs = {}
    for ns in (4, 6, 7):
        m.set_precision(ns)
        errs[ns] = rel_l2(m.tendency(BarotropicState(z)).cpu().numpy(), ref)
    m.set_precision(0)
    assert rel_l2(m.tendency(BarotropicState(z)).cpu().numpy(), ref) == 0.0
    print("split points tendency rel errors", errs, flush=True)
    assert errs[6] < 1e-11 and errs[7] < 1e-13
    assert 1e-12 < errs[4] < 1e-6


def test_split_points_swe_tendency_matches_fp64():
    import torch
    from aeolus_amd.physics import ShallowWaterModel
    from aeolus_amd.run.initial_conditions import swe_spectra
    print("case split points swe L=21 res=4", flush=True)
    m = ShallowWaterModel(_planet(21, 4), gravity=9.80616, mean_depth=1e4)
    y = torch.as_tensor(swe_spectra(21, R, 9.80616, seed=3), device="cuda")
    ref = m.tendency(y).cpu().numpy()
    m.set_precision(6)
    got = m.tendency(y).cpu().numpy()
    errs = [rel_l2(got[f], ref[f]) for f in range(3)]
    print("split points swe rel errors", errs, flush=True)
    for f in range(3):
        assert errs[f] < 1e-11, f


def test_split_points_c2_trajectory_vs_golden():
    """C2 (geodesic res 4, L = 21, 96 x 900 s) with 6-slice Y-basis
    contractions against the FP64 oracle's fixture."""
    import torch
    from aeolus_amd.physics import BarotropicVorticity
    from aeolus_amd.run.initial_conditions import spectrum_l15
    print("case split points golden c2", flush=True)
    g = np.load(os.path.join(GOLDEN, "c2.npz"))
    z0 = spectrum_l15(21, np.random.default_rng(int(g["seed"])), rms=5e-5)
    model = BarotropicVorticity(_planet(21, 4))
    model.set_precision(6)
    zT = model.rk4(torch.as_tensor(z0, device="cuda"), float(g["dt"]), int(g["nsteps"]))
    li, mi = np.tril_indices(22)
    err = rel_l2(zT.cpu().numpy()[li, mi], g["zT"])
    print(f"split(6) points c2 vs oracle: {err:.3e}", flush=True)
    assert err < 1e-11


def test_split_points_rejects_mismatched_tables():
    from aeolus_amd.physics import BarotropicVorticity
    pl = _planet(10, 3)
    m = BarotropicVorticity(pl)
    m.set_precision(6)
    m2 = BarotropicVorticity(pl)
    with pytest.raises(Exception):
        # the plan holds 6-slice tables; a model cannot run 5 without rebuilding them
        from aeolus_amd import _native as nat
        nat.check(nat.lib().aeo_model_set_precision(m2._model, 5), "set_precision")
