"""Split-precision (INT8 radix-254 digit MFMA, exact INT32 accumulation,
FP64 recombination) Legendre contractions against the FP64 path and the
oracle.

Bars (measured on MI355X at L = 170: 4 slices 1.1e-7, 6 slices 2.9e-12,
7 slices 2.2e-15 relative L2): 6 slices within 5e-12 of the FP64 tendency,
7 slices within 1e-13 (FP64 level), 4 slices coarser than 1e-12 but within
5e-7 — the digits carry what the slice count says (254^-4 = 2.4e-10 per
entry before the sums).
"""
import math
import warnings

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

R = 6.371e6
OM = 2.0 * math.pi / 86400.0


def _planet(L, nlat, nlon):
    from aeolus_amd.planet import Planet
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return Planet.with_radius(R, OM, grid_type="latlon", nlat=nlat, nlon=nlon, l_max=L,
                                  product_quadrature="fine")


@pytest.mark.parametrize("L,nlat,nlon", [(21, 32, 64), (85, 128, 256), (170, 256, 512),
                                         (200, 320, 640)])
def test_split_bve_tendency_matches_fp64(L, nlat, nlon):
    import torch
    from aeolus_amd.physics import BarotropicState, BarotropicVorticity
    from aeolus_amd.run.initial_conditions import spectrum_l15
    print(f"case split bve L={L}", flush=True)
    pl = _planet(L, nlat, nlon)
    m = BarotropicVorticity(pl)
    z = torch.as_tensor(spectrum_l15(L, np.random.default_rng(L), rms=5e-5), device="cuda")
    ref = m.tendency(BarotropicState(z)).cpu().numpy()
    errs = {}
    for ns in (4, 6, 7):
        m.set_precision(ns)
        got = m.tendency(BarotropicState(z)).cpu().numpy()
        errs[ns] = rel_l2(got, ref)
    m.set_precision(0)
    print("split tendency rel errors", errs)
    assert errs[6] < 5e-12 and errs[7] < 1e-13
    assert 1e-12 < errs[4] < 5e-7


def test_split_batched_members_match_single():
    import torch
    from aeolus_amd.physics.barotropic import BarotropicBatch
    from aeolus_amd.physics import BarotropicVorticity
    from aeolus_amd.run.initial_conditions import spectrum_l15
    L = 42
    pl = _planet(L, 64, 128)
    zs = np.stack([spectrum_l15(L, np.random.default_rng(7 + k), rms=5e-5) for k in range(6)])
    single = BarotropicVorticity(pl)
    single.set_precision(6)
    for mb in (1, 2, 4):
        batch = BarotropicBatch(pl, 6)
        batch.set_precision(6)
        batch.set_tiles(mb)
        out = batch.rk4(torch.as_tensor(zs, device="cuda"), 600.0, 3)
        for k in range(6):
            ref = single.rk4(torch.as_tensor(zs[k], device="cuda"), 600.0, 3)
            assert torch.equal(out[k], ref), (mb, k)


@pytest.mark.parametrize("L,nlat,nlon,nmem,ns", [(42, 64, 128, 3, 7), (85, 128, 256, 40, 6),
                                                 (170, 256, 512, 20, 5), (200, 320, 640, 5, 6)])
def test_split_gemm_form_matches_fused_bitwise(L, nlat, nlon, nmem, ns):
    """The member-batched GEMM form (tiles 0: analysis, epilogue and synthesis
    kernels spanning the batch) against the fused stage kernel at the same
    slice count: the same digits, exact INT32 sums and FP64 recombination, so
    RK4 steps and tendencies are bitwise equal (ragged member groups
    included: 3, 40 and 20 members)."""
    import torch
    from aeolus_amd.physics import BarotropicState
    from aeolus_amd.physics.barotropic import BarotropicBatch
    from aeolus_amd.run.initial_conditions import spectrum_l15
    print(f"case split gemm form L={L} nmem={nmem} ns={ns}", flush=True)
    pl = _planet(L, nlat, nlon)
    zs = torch.as_tensor(np.stack([spectrum_l15(L, np.random.default_rng(50 + k), rms=5e-5)
                                   for k in range(nmem)]), device="cuda")
    fused = BarotropicBatch(pl, nmem, viscosity=1e5)
    fused.set_precision(ns)
    fused.set_tiles(4)
    gemm = BarotropicBatch(pl, nmem, viscosity=1e5)
    gemm.set_precision(ns)
    gemm.set_tiles(0)
    a = fused.rk4(zs, 600.0, 3)
    b = gemm.rk4(zs, 600.0, 3)
    assert torch.equal(a, b)
    for tiles in ((128, 4, 4), (64, 4, 2), (128, 8, 2)):
        gemm.set_gemm_tiles(*tiles)
        assert torch.equal(gemm.rk4(zs, 600.0, 3), a), tiles
    gemm.set_gemm_tiles(64, 8, 4)
    ka = fused.tendency(BarotropicState(zs))
    kb = gemm.tendency(BarotropicState(zs))
    assert torch.equal(ka, kb)
    # and the FP64 path within the slice count's accuracy
    gemm.set_precision(0)
    with pytest.raises(RuntimeError, match="split precision"):
        gemm.rk4(zs, 600.0, 1)
    gemm.set_precision(ns)
    ref = BarotropicBatch(pl, nmem, viscosity=1e5)
    r = ref.rk4(zs, 600.0, 3)
    err = rel_l2(b.cpu().numpy(), r.cpu().numpy())
    print(f"  gemm form vs fp64 rel-L2 {err:.3e}", flush=True)
    assert err < {5: 1e-9, 6: 5e-12, 7: 1e-13}[ns]


def test_split_gemm_form_refused_where_unsupported():
    from aeolus_amd.physics import ShallowWaterModel
    print("case split gemm form refusals", flush=True)
    pl = _planet(21, 32, 64)
    swe = ShallowWaterModel(pl, gravity=9.80616, mean_depth=1e4)
    with pytest.raises(ValueError, match="GEMM"):
        swe.set_tiles(0)
    from aeolus_amd.physics.barotropic import BarotropicBatch
    b = BarotropicBatch(pl, 2)
    with pytest.raises(ValueError, match="syn_cols"):
        b.set_gemm_tiles(96, 8)
    with pytest.raises(ValueError, match="epi_members"):
        b.set_gemm_tiles(64, 2)
    with pytest.raises(ValueError, match="syn_ntiles"):
        b.set_gemm_tiles(64, 8, 3)


def test_split_swe_tendency_matches_fp64():
    import torch
    from aeolus_amd.physics import ShallowWaterModel
    from aeolus_amd.run.initial_conditions import swe_spectra
    L = 85
    pl = _planet(L, 128, 256)
    m = ShallowWaterModel(pl, mean_depth=1e4)
    y = torch.as_tensor(swe_spectra(L, R, 9.80616, seed=3), device="cuda")
    ref = m.tendency(y).cpu().numpy()
    m.set_precision(6)
    got = m.tendency(y).cpu().numpy()
    for f in range(3):
        assert rel_l2(got[f], ref[f]) < 1e-12, f


def test_split_bve_day_vs_oracle_l85():
    """One simulated day at L = 85 (288 x 300 s): split (6 slices) vs the FP64
    oracle restatement; the error is reported and bounded."""
    import torch
    from aeolus_amd.physics import BarotropicVorticity
    from aeolus_amd.run.initial_conditions import spectrum_l15
    L = 85
    pl = _planet(L, 128, 256)
    m = BarotropicVorticity(pl)
    z0 = spectrum_l15(L, np.random.default_rng(85), rms=5e-5)
    fp64 = m.rk4(torch.as_tensor(z0, device="cuda"), 300.0, 288).cpu().numpy()
    m.set_precision(6)
    sp = m.rk4(torch.as_tensor(z0, device="cuda"), 300.0, 288).cpu().numpy()
    err = rel_l2(sp, fp64)
    print(f"split(6) vs fp64 after 1 day at L=85: {err:.3e}")
    assert err < 1e-11


@pytest.mark.parametrize("L,nlat,nlon", [(20, 30, 60), (33, 51, 100)])
def test_split_odd_grids_match_fp64(L, nlat, nlon):
    """Odd latitude counts (an equator row, nh not a multiple of 16) and
    non-power-of-two longitudes: padding rows and columns of the digit tiles."""
    import torch
    from aeolus_amd.physics import BarotropicState, BarotropicVorticity
    from aeolus_amd.run.initial_conditions import spectrum_l15
    print(f"case split odd grid L={L} {nlat}x{nlon}", flush=True)
    pl = _planet(L, nlat, nlon)
    m = BarotropicVorticity(pl, viscosity=1e5)
    z = torch.as_tensor(spectrum_l15(L, np.random.default_rng(L + 1), rms=5e-5), device="cuda")
    ref = m.tendency(BarotropicState(z)).cpu().numpy()
    y_ref = m.rk4(z, 600.0, 20).cpu().numpy()
    m.set_precision(6)
    assert rel_l2(m.tendency(BarotropicState(z)).cpu().numpy(), ref) < 5e-12
    assert rel_l2(m.rk4(z, 600.0, 20).cpu().numpy(), y_ref) < 1e-11
