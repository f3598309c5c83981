"""HIP primitive-equation tendency and weak-form vector curl/divergence
against the CPU oracle, plus the exact properties the reference tests.

Bars: vector analysis within 1e-12 of the oracle's scale (and of the closed
forms curl = lap psi, div = lap chi on the Gauss 3/2-rule grid); the PE
tendency within 1e-11 of each row block's scale (zeta, delta, T, ln p_s) on
the Gauss (L = 15, 32 x 64; L = 42, 64 x 128 with 10 non-uniform levels) and
geodesic (L = 10, res 3) product samplings; at the measured Gauss L = 85
configuration, exact rest and BVE degeneracy;
exact rest bitwise zero; BVE degeneracy within 1e-12 of the barotropic
tendency's scale; zero flow with structured T gives -lap(Phi) within 1e-12;
state-grid wind, temperature extrema and characteristic speed within 1e-12 of
the oracle (Lamb bound at rest); one validated RK4 step within 1e-11 of the
oracle's increment.
"""
import math
import warnings

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

T0, PS0 = 280.0, 1.0e5
SQRT4PI = math.sqrt(4.0 * math.pi)
NONUNIFORM = (0.0, 0.05, 0.15, 0.30, 0.50, 0.75, 1.0)
CASES = {"latlon": dict(L=15, nlat=32, nlon=64, res=0), "geodesic": dict(L=10, nlat=0, nlon=0,
                                                                          res=3)}
_cache = {}


def _planets(kind):
    if kind not in _cache:
        from aeolus_amd.planet import Planet
        from oracle.models import Planet as OPlanet, planet_constants
        c = CASES[kind]
        R, om = planet_constants(24.0)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            pl = Planet.with_radius(R, om, grid_type=kind, nlat=c["nlat"], nlon=c["nlon"],
                                    grid_resolution=c["res"], l_max=c["L"],
                                    product_quadrature="fine")
        op = OPlanet(l_max=c["L"], radius=R, omega=om, grid_type=kind, nlat=c["nlat"],
                     nlon=c["nlon"], resolution=c["res"])
        _cache[kind] = (pl, op)
    return _cache[kind]


def _np(t):
    return t.detach().cpu().numpy()


def _band_limited(L, K, seed):
    from oracle import pe
    y = pe.isothermal_rest_state(L, K, T0, PS0)
    rng = np.random.default_rng(seed)
    n, cut = L + 1, (2 * L) // 3

    def pert(amp, zero_mono):
        p = amp * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        p[:, 0] = p[:, 0].real
        p *= np.tril(np.ones((n, n)))
        p[cut + 1:, :] = 0.0
        if zero_mono:
            p[0, 0] = 0.0
        return p
    for k in range(K):
        y[k] += pert(2e-6, True)
        y[K + k] += pert(2e-6, True)
        y[2 * K + k] += pert(0.5, False)
    y[3 * K] += pert(1e-3, True)
    return y


def _potentials(L, seed, top):
    rng = np.random.default_rng(seed)
    a = (rng.standard_normal((L + 1, L + 1)) + 1j * rng.standard_normal((L + 1, L + 1))) * 1e6
    a = np.tril(a)
    a[:, 0] = a[:, 0].real
    a[0, :] = 0.0
    a[top + 1:, :] = 0.0
    return a


def _vector(op, psi, chi):
    so, psh, c = op.so, op.psh, op.pcoslat
    p_lam = psh.inv_transform(so.d_lambda(psi))
    p_snt = psh.inv_transform(so.sin_theta_d_theta(psi)) / so.R
    c_lam = psh.inv_transform(so.d_lambda(chi))
    c_snt = psh.inv_transform(so.sin_theta_d_theta(chi)) / so.R
    return (p_snt + c_lam) / c, (p_lam - c_snt) / c


@pytest.mark.parametrize("kind", ["latlon", "geodesic"])
def test_vector_curl_div_matches_oracle(kind):
    import torch
    from oracle import pe
    print(f"case vector curl/div {kind}", flush=True)
    pl, op = _planets(kind)
    L = op.l_max
    psi, chi = _potentials(L, 1, L - 1), _potentials(L, 2, L - 1)
    fe, fn = _vector(op, psi, chi)
    ext = pe.extended_transform(op)
    for trunc in (False, True):
        rc, rd = pe.vector_curl_div(op, ext, fe, fn, truncate=trunc)
        gc, gd = pl.so.vector_curl_div_spectral(torch.as_tensor(fe, device="cuda"),
                                                torch.as_tensor(fn, device="cuda"),
                                                truncate=trunc)
        scale = max(np.abs(rc).max(), np.abs(rd).max())
        ec, ed = np.abs(_np(gc) - rc).max() / scale, np.abs(_np(gd) - rd).max() / scale
        print(f"  truncate={trunc}: curl {ec:.2e} div {ed:.2e}", flush=True)
        assert ec < 1e-12 and ed < 1e-12
    if kind == "latlon":
        gc, gd = pl.so.vector_curl_div_spectral(torch.as_tensor(fe, device="cuda"),
                                                torch.as_tensor(fn, device="cuda"), truncate=False)
        lap = op.so.lap_eigs[:, None]
        scale = max(np.abs(lap * psi).max(), np.abs(lap * chi).max())
        assert np.abs(_np(gc) - lap * psi).max() < 1e-12 * scale
        assert np.abs(_np(gd) - lap * chi).max() < 1e-12 * scale
    # batched call: each field as alone
    fe2 = torch.as_tensor(np.stack([fe, 2.0 * fn]), device="cuda")
    fn2 = torch.as_tensor(np.stack([fn, -fe]), device="cuda")
    bc, bd = pl.so.vector_curl_div_spectral(fe2, fn2)
    for f in range(2):
        sc, sd = pl.so.vector_curl_div_spectral(fe2[f], fn2[f])
        assert torch.allclose(bc[f], sc, rtol=0, atol=1e-14 * float(sc.abs().max()))
        assert torch.allclose(bd[f], sd, rtol=0, atol=1e-14 * float(sd.abs().max()))


def test_vector_curl_div_zero_field_is_bitwise_zero():
    import torch
    pl, _ = _planets("latlon")
    z = torch.zeros(pl.so.extended_product_transform().n_points, dtype=torch.float64,
                    device="cuda")
    c, d = pl.so.vector_curl_div_spectral(z, z)
    assert float(c.abs().max()) == 0.0 and float(d.abs().max()) == 0.0


@pytest.mark.parametrize("kind,sigma", [("latlon", 5), ("geodesic", 5), ("latlon", "nonuniform"),
                                        ("latlon", 1), ("latlon", 32), ("geodesic", 32)])
def test_pe_tendency_matches_oracle(kind, sigma):
    import torch
    from aeolus_amd.physics import PrimitiveEquationsModel, SigmaGrid
    from oracle import pe
    print(f"case pe tendency {kind} sigma={sigma}", flush=True)
    pl, op = _planets(kind)
    L = op.l_max
    interfaces = NONUNIFORM if sigma == "nonuniform" else tuple(
        k / sigma for k in range(sigma + 1))
    K = len(interfaces) - 1
    phi_s = None
    if sigma == "nonuniform":
        # band-limited terrain at the 2/3 cut
        phi_s = _potentials(L, 9, (2 * L) // 3) * 1e-3
        phi_s[:, (2 * L) // 3 + 1:] = 0.0
        phi_s[0, 0] = 4000.0
    m = PrimitiveEquationsModel(pl, SigmaGrid(interfaces), surface_geopotential_lm=phi_s)
    om = pe.PEModel(op, pe.Sigma(interfaces), phi_surface_lm=phi_s)
    y = _band_limited(L, K, 3)
    got = _np(m.tendency(torch.as_tensor(y, device="cuda")))
    ref = pe.pe_tendency(om, y)
    blocks = {"zeta": slice(0, K), "delta": slice(K, 2 * K), "T": slice(2 * K, 3 * K),
              "lnps": slice(3 * K, 3 * K + 1)}
    errs = {b: np.abs(got[s] - ref[s]).max() / np.abs(ref[s]).max() for b, s in blocks.items()}
    print("  block rel errors", {b: f"{e:.2e}" for b, e in errs.items()}, flush=True)
    for b, e in errs.items():
        assert e < 1e-11, b
    assert not got[:2 * K, 0, :].any()


@pytest.mark.parametrize("kind", ["latlon", "geodesic"])
def test_pe_exact_rest_is_bitwise_zero(kind):
    import torch
    from aeolus_amd.physics import PrimitiveEquationsModel, SigmaGrid, isothermal_rest_state
    from aeolus_amd.run.engine import rk4_step_array
    pl, op = _planets(kind)
    m = PrimitiveEquationsModel(pl, SigmaGrid.uniform(5))
    st = isothermal_rest_state(op.l_max, 5, temperature=T0, surface_pressure=PS0, device="cuda")
    out = m.tendency(st.coeffs)
    assert tuple(out.shape) == tuple(st.coeffs.shape)
    assert float(out.abs().max()) == 0.0
    y1 = rk4_step_array(m.tendency, st.coeffs, 0.0, 600.0)
    assert torch.equal(y1, st.coeffs)


@pytest.mark.parametrize("kind", ["latlon", "geodesic"])
def test_pe_bve_degeneracy(kind):
    """delta = 0, per-level uniform T, uniform ln p_s: each level's zeta
    tendency is the barotropic model's (the fused BVE kernels)."""
    import torch
    from aeolus_amd.physics import (BarotropicState, BarotropicVorticity,
                                    PrimitiveEquationsModel, SigmaGrid)
    from oracle import pe
    pl, op = _planets(kind)
    L = op.l_max
    z = _potentials(L, 5, L) * 1e-11
    m = PrimitiveEquationsModel(pl, SigmaGrid.uniform(4))
    y = pe.isothermal_rest_state(L, 4, T0, PS0)
    for k in range(4):
        y[k] += z
        y[8 + k, 0, 0] = (T0 - 6.0 * k) * SQRT4PI
    out = _np(m.tendency(torch.as_tensor(y, device="cuda")))
    bve = _np(BarotropicVorticity(pl, viscosity=0.0).tendency(
        BarotropicState(torch.as_tensor(z, device="cuda"))))
    scale = np.abs(bve).max()
    errs = [np.abs(out[k] - bve).max() / scale for k in range(4)]
    print(f"case pe bve degeneracy {kind}: {max(errs):.2e}", flush=True)
    assert max(errs) <= 1e-12


def test_pe_zero_flow_structured_temperature():
    import torch
    from aeolus_amd.physics import PrimitiveEquationsModel, SigmaGrid
    from oracle import pe
    pl, op = _planets("latlon")
    m = PrimitiveEquationsModel(pl, SigmaGrid.uniform(5))
    y = pe.isothermal_rest_state(15, 5, T0, PS0)
    for k in range(5):
        y[10 + k, 0, 0] = (T0 - 7.0 * k) * SQRT4PI
        y[10 + k, 4, 2] = 3.0 - 1.0j
        y[10 + k, 2, 0] = 1.5 * (k + 1)
    out = _np(m.tendency(torch.as_tensor(y, device="cuda")))
    assert not out[:5].any()
    phi, _ = pe.hydrostatic(pe.Sigma.uniform(5), y[10:15], np.zeros((16, 16), complex), pe.R_DRY)
    expected = -op.so.lap_eigs[None, :, None] * phi
    scale = np.abs(expected).max()
    assert scale > 0 and np.abs(out[5:10] - expected).max() < 1e-12 * scale


def test_pe_rejects_bad_inputs():
    import torch
    from aeolus_amd.physics import (PrimitiveEquationsModel, PrimitiveEquationsStateError,
                                    SigmaGrid)
    pl, _ = _planets("latlon")
    with pytest.raises(ValueError):
        PrimitiveEquationsModel(pl, SigmaGrid.uniform(3), r_dry=-1.0)
    with pytest.raises(ValueError):
        PrimitiveEquationsModel(pl, SigmaGrid.uniform(3), cp_dry=100.0)
    high = np.zeros((16, 16), complex)
    high[14, 3] = 1.0
    with pytest.raises(ValueError):
        PrimitiveEquationsModel(pl, SigmaGrid.uniform(3), surface_geopotential_lm=high)
    m = PrimitiveEquationsModel(pl, SigmaGrid.uniform(3))
    with pytest.raises(PrimitiveEquationsStateError):
        m.tendency(torch.zeros((13, 16, 16), dtype=torch.complex128, device="cuda"))


@pytest.mark.parametrize("kind", ["latlon", "geodesic"])
def test_pe_state_wind_extrema_and_speed_match_oracle(kind):
    import torch
    from aeolus_amd.physics import PrimitiveEquationsModel, PrimitiveEquationsState, SigmaGrid
    from oracle import pe
    pl, op = _planets(kind)
    L = op.l_max
    m = PrimitiveEquationsModel(pl, SigmaGrid.uniform(5))
    om = pe.PEModel(op, pe.Sigma.uniform(5))
    y = _band_limited(L, 5, 11)
    st = PrimitiveEquationsState(torch.as_tensor(y, device="cuda"))
    u, v = m.wind_on_state_grid(st)
    ru, rv = pe.state_wind(om, y)
    scale = max(np.abs(ru).max(), np.abs(rv).max())
    assert np.abs(_np(u) - ru).max() < 1e-12 * scale
    assert np.abs(_np(v) - rv).max() < 1e-12 * scale
    lo, hi = m.temperature_extrema(st)
    rlo, rhi = pe.temperature_extrema(om, y)
    assert lo == pytest.approx(rlo, rel=1e-12) and hi == pytest.approx(rhi, rel=1e-12)
    assert m.max_characteristic_speed(st) == pytest.approx(pe.max_characteristic_speed(om, y),
                                                           rel=1e-12)
    rest = pe.isothermal_rest_state(L, 5, T0, PS0)
    lamb = math.sqrt(m.gamma * m.r_dry * T0)
    assert m.max_characteristic_speed(
        PrimitiveEquationsState(torch.as_tensor(rest, device="cuda"))) == pytest.approx(
        lamb, rel=1e-9)
    m.validate_state(st)


def test_pe_validation_rejects_invalid_states():
    """The reference's hard failures (``tests/test_primitive_equations.py:446-492``)."""
    import torch
    from aeolus_amd.physics import (PrimitiveEquationsModel, PrimitiveEquationsState,
                                    PrimitiveEquationsStateError, SigmaGrid,
                                    isothermal_rest_state)
    pl, _ = _planets("latlon")
    m = PrimitiveEquationsModel(pl, SigmaGrid.uniform(5))

    def rest():
        return isothermal_rest_state(15, 5, temperature=T0, surface_pressure=PS0, device="cuda")
    m.validate_state(rest())
    cases = []
    b = rest(); b.temperature[1, 0, 0] = -50.0 * SQRT4PI; cases.append((b, "temperature"))
    b = rest(); b.temperature[0, 0, 0] = 0.0; cases.append((b, "temperature"))
    b = rest(); b.coeffs[0, 3, 2] = math.nan; cases.append((b, "NaN/Inf"))
    b = rest(); b.zeta[2, 0, 0] = 1e-3; cases.append((b, "zeta monopole .* level 3"))
    b = rest(); b.delta[0, 0, 0] = 1e-3; cases.append((b, "delta monopole"))
    b = rest(); b.ln_ps[0, 0] = 1000.0 * SQRT4PI; cases.append((b, "surface pressure"))
    cases.append((PrimitiveEquationsState.zeros(15, 6, device="cuda"), "shape"))
    for st, msg in cases:
        with pytest.raises(PrimitiveEquationsStateError, match=msg):
            m.validate_state(st)


def test_pe_rk4_step_with_stage_validation_matches_oracle():
    """One RK4 step through rk4_step_array with validate_state on every stage
    (the reference's run_pe step) against the oracle's RK4 of its tendency."""
    import torch
    from aeolus_amd.physics import PrimitiveEquationsModel, PrimitiveEquationsState, SigmaGrid
    from aeolus_amd.run.engine import rk4_step_array
    from oracle import pe
    from oracle.models import rk4_step_array as ork4
    pl, op = _planets("latlon")
    m = PrimitiveEquationsModel(pl, SigmaGrid.uniform(5))
    om = pe.PEModel(op, pe.Sigma.uniform(5))
    y = _band_limited(15, 5, 13)
    seen = []

    def validator(ys):
        seen.append(1)
        m.validate_state(PrimitiveEquationsState(ys))
    got = _np(rk4_step_array(m.tendency, torch.as_tensor(y, device="cuda"), 0.0, 300.0,
                             stage_validator=validator))
    ref = ork4(lambda s: pe.pe_tendency(om, s), y, 300.0)
    assert len(seen) == 3
    for s in (slice(0, 5), slice(5, 10), slice(10, 15), slice(15, 16)):
        # 1e-11 of the step's increment, plus a few ulps of the state it is added to
        d = ref[s] - y[s]
        assert np.abs(got[s] - ref[s]).max() < 1e-11 * np.abs(d).max() + 4e-16 * np.abs(y[s]).max()


def test_run_pe_matches_oracle(tmp_path):
    """run_pe (fixed 300 s step, count snapshots, stage validation) against
    the oracle stepped with the same dt sequence: snapshots within 1e-12 of
    each row block's scale, every diagnostics row within 1e-10."""
    import csv
    import torch
    from aeolus_amd.physics import PrimitiveEquationsModel, PrimitiveEquationsState, SigmaGrid
    from aeolus_amd.run import IntegrationScheduler, run_pe
    from oracle import pe
    from oracle.diagnostics import pe_record_row
    from oracle.models import rk4_step_array as ork4
    print("case run_pe latlon L=15 K=4", flush=True)
    pl, op = _planets("latlon")
    m = PrimitiveEquationsModel(pl, SigmaGrid.uniform(4))
    om = pe.PEModel(op, pe.Sigma.uniform(4))
    y0 = _band_limited(15, 4, 17)
    times = [0.0, 1200.0, 2700.0]
    run_pe(m, PrimitiveEquationsState(torch.as_tensor(y0, device="cuda")), dt_seconds=300.0,
           t_end_days=0.03125, out_dir=tmp_path, snapshot_times=times)
    got = np.load(tmp_path / "pe_coeffs.npy")
    assert np.array_equal(np.load(tmp_path / "pe_snapshot_times.npy"), np.asarray(times))
    rows = list(csv.DictReader(open(tmp_path / "diagnostics" / "timeseries.csv")))
    sched = IntegrationScheduler(2700.0, mode="count", snapshot_times=times)
    y, snaps, ref_rows = y0, [], [pe_record_row(om, y0, 0.0, pl.grid.cfl_length_scale)]
    while (ev := sched.next_event(300.0)) is not None:
        if ev[0] == "store":
            snaps.append(y)
            continue
        y = ork4(lambda s: pe.pe_tendency(om, s), y, ev[1])
        ref_rows.append(pe_record_row(om, y, ev[1], pl.grid.cfl_length_scale,
                                      mass0=ref_rows[0]["total_mass"]))
    assert got.shape == (3, 13, 16, 16) and len(rows) == len(ref_rows) == 10
    for i, s in enumerate(snaps):
        for b in (slice(0, 4), slice(4, 8), slice(8, 12), slice(12, 13)):
            assert np.abs(got[i][b] - s[b]).max() <= 1e-12 * np.abs(s[b]).max(), (i, b)
    for r, rr in zip(rows, ref_rows):
        for key, val in rr.items():
            if key == "mass_rel_drift":
                assert abs(float(r[key]) - val) < 1e-13, key
            elif key == "courant" and float(r["dt_s"]) == 0.0:
                assert math.isnan(float(r[key]))
            else:
                assert float(r[key]) == pytest.approx(val, rel=1e-10, abs=1e-300), key


def _terrain_model(kind):
    from aeolus_amd.physics import PrimitiveEquationsModel, SigmaGrid, Topography
    pl, _ = _planets(kind)
    L = pl.sh.l_max
    topo = Topography.mountain(pl, height_m=1500.0, lat_deg=25.0, lon_deg=60.0, width_deg=25.0,
                               l_cut=(2 * L) // 3)
    phi_s = topo.surface_geopotential_lm(9.80616)
    return PrimitiveEquationsModel(pl, SigmaGrid.uniform(4), surface_geopotential_lm=phi_s)


@pytest.mark.parametrize("kind,rtol", [("latlon", 1e-13), ("geodesic", 1e-2)])
def test_pe_orographic_rest_is_balanced(kind, rtol):
    """Isothermal rest over band-limited terrain: the pressure-gradient terms
    cancel (``tests/test_pe_orographic_rest.py:471-502``): momentum blocks
    within rtol of max|lap Phi_s|, T and ln p_s blocks bitwise zero."""
    from aeolus_amd.run.initial_conditions import make_pe_ic
    m = _terrain_model(kind)
    st = make_pe_ic("orographic_isothermal_rest", m, temperature=T0, surface_pressure=PS0)
    out = m.tendency(st.coeffs)
    scale = float((m.lap_eigs[:, None] * m.phi_surface_lm).abs().max())
    K = m.nlev
    res = {"zeta": float(out[:K].abs().max()), "delta": float(out[K:2 * K].abs().max()),
           "T": float(out[2 * K:3 * K].abs().max()), "lnps": float(out[3 * K].abs().max())}
    print(f"case orographic rest {kind}: scale {scale:.3e}", res, flush=True)
    assert scale > 0 and res["T"] == 0.0 and res["lnps"] == 0.0
    assert res["zeta"] <= rtol * scale and res["delta"] <= rtol * scale


def test_pe_orographic_rest_preserved_by_run(tmp_path):
    """Five fixed 300 s steps through run_pe keep the Gauss orographic rest:
    max wind <= 1e-12 m/s, mass drift <= 1e-14 (``tests/test_pe_orographic_rest.py:84-101``)."""
    import csv
    from aeolus_amd.run import run_pe
    from aeolus_amd.run.initial_conditions import make_pe_ic
    m = _terrain_model("latlon")
    st = make_pe_ic("orographic_isothermal_rest", m, temperature=T0, surface_pressure=PS0)
    run_pe(m, st, dt_seconds=300.0, t_end_days=1500.0 / 86400.0, out_dir=tmp_path,
           snapshot_times=[0.0])
    rows = list(csv.DictReader(open(tmp_path / "diagnostics" / "timeseries.csv")))
    assert len(rows) == 6
    for r in rows:
        assert float(r["max_wind_ms"]) <= 1e-12
        assert abs(float(r["mass_rel_drift"])) <= 1e-14


def test_pe_thermal_wave_preset():
    import torch
    from aeolus_amd.physics import PrimitiveEquationsModel, SigmaGrid
    from aeolus_amd.run.initial_conditions import make_pe_ic
    pl, _ = _planets("latlon")
    m = PrimitiveEquationsModel(pl, SigmaGrid.uniform(3))
    st = make_pe_ic("thermal_wave", m, temperature=260.0, surface_pressure=101325.0,
                    thermal_amplitude=1.0)
    assert torch.all(st.temperature[:, 2, 2] == 1.0)
    m.validate_state(st)
    out = m.tendency(st.coeffs)
    assert float(out[3:6].abs().max()) > 0.0          # the unbalanced T drives divergence
    with pytest.raises(ValueError):
        make_pe_ic("nope", m, temperature=260.0, surface_pressure=101325.0)


def test_pe_create_rejects_mismatched_plans():
    """aeo_pe_create checks that the extended plan is degree L + 1 on the
    product plan's points; aeo_vector_curl_div checks its arguments."""
    import ctypes
    import torch
    from aeolus_amd import _native as nat
    pl, _ = _planets("latlon")
    so = pl.so
    prod = so.product_transform(15, max_fields=8)
    interfaces = np.asarray([0.0, 0.5, 1.0])
    h = ctypes.c_void_p()
    for ext in (so.product_transform(17), so.product_transform(15)):
        rc = nat.lib().aeo_pe_create(prod._plan, ext._plan, 2, nat.host_ptr(interfaces), 6.4e6,
                                     7.29e-5, 287.04, 1004.64, None, ctypes.byref(h))
        assert rc == nat.AEO_ERR_INVALID and not h.value
    ext = so.product_transform(16)
    for bad in (np.asarray([0.1, 0.5, 1.0]), np.asarray([0.0, 0.5, 0.9]),
                np.asarray([0.0, 0.6, 0.6])):
        rc = nat.lib().aeo_pe_create(prod._plan, ext._plan, 2, nat.host_ptr(bad), 6.4e6,
                                     7.29e-5, 287.04, 1004.64, None, ctypes.byref(h))
        assert rc == nat.AEO_ERR_INVALID and not h.value
    rc = nat.lib().aeo_pe_create(prod._plan, ext._plan, 2, nat.host_ptr(interfaces), 6.4e6,
                                 7.29e-5, 287.04, 200.0, None, ctypes.byref(h))
    assert rc == nat.AEO_ERR_INVALID
    z = torch.zeros(ext.n_points, dtype=torch.float64, device="cuda")
    out = torch.empty((16, 16), dtype=torch.complex128, device="cuda")
    rc = nat.lib().aeo_vector_curl_div(ext._plan, -1.0, z.data_ptr(), z.data_ptr(), 1, 1,
                                       out.data_ptr(), out.data_ptr(), z.data_ptr(), nat.stream())
    assert rc == nat.AEO_ERR_INVALID
    with pytest.raises(ValueError):
        so.vector_curl_div_spectral(z[:-1], z[:-1])


def test_pe_tendency_larger_grid_matches_oracle():
    """Gauss 64 x 128 (3/2-rule product 64 x 128), L = 42, 10 non-uniform
    levels against the oracle: each row block within 1e-11 of its scale."""
    import torch
    from aeolus_amd.physics import PrimitiveEquationsModel, SigmaGrid
    from aeolus_amd.planet import Planet
    from oracle import pe
    from oracle.models import Planet as OPlanet, planet_constants
    print("case pe tendency latlon L=42 K=10", flush=True)
    R, om = planet_constants(24.0)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        pl = Planet.with_radius(R, om, grid_type="latlon", nlat=64, nlon=128, l_max=42,
                                product_quadrature="fine")
    op = OPlanet(l_max=42, radius=R, omega=om, grid_type="latlon", nlat=64, nlon=128)
    s = np.concatenate([[0.0], np.cumsum(np.linspace(1.0, 2.0, 10))])
    interfaces = tuple(s / s[-1])
    interfaces = interfaces[:-1] + (1.0,)
    m = PrimitiveEquationsModel(pl, SigmaGrid(interfaces))
    om_ = pe.PEModel(op, pe.Sigma(interfaces))
    y = _band_limited(42, 10, 23)
    got = _np(m.tendency(torch.as_tensor(y, device="cuda")))
    ref = pe.pe_tendency(om_, y)
    for b, sl in (("zeta", slice(0, 10)), ("delta", slice(10, 20)), ("T", slice(20, 30)),
                  ("lnps", slice(30, 31))):
        e = np.abs(got[sl] - ref[sl]).max() / np.abs(ref[sl]).max()
        print(f"  {b}: {e:.2e}", flush=True)
        assert e < 1e-11, b


@pytest.mark.parametrize("L,nlat,nlon", [(85, 128, 256), (170, 256, 512)])
def test_pe_properties_at_benchmark_size(L, nlat, nlon):
    """Gauss 128 x 256 / L = 85 (the measured PE configuration) and
    256 x 512 / L = 170, 8 levels: exact rest is bitwise zero and the
    BVE-degenerate state reproduces the fused barotropic kernels' tendency
    within 1e-12 of its scale."""
    import torch
    from aeolus_amd.physics import (BarotropicState, BarotropicVorticity,
                                    PrimitiveEquationsModel, SigmaGrid, isothermal_rest_state)
    from aeolus_amd.planet import Planet
    from oracle.models import planet_constants
    print(f"case pe properties latlon L={L} K=8", flush=True)
    R, om = planet_constants(24.0)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        pl = Planet.with_radius(R, om, grid_type="latlon", nlat=nlat, nlon=nlon, l_max=L,
                                product_quadrature="fine")
    m = PrimitiveEquationsModel(pl, SigmaGrid.uniform(8))
    st = isothermal_rest_state(L, 8, temperature=T0, surface_pressure=PS0, device="cuda")
    assert float(m.tendency(st.coeffs).abs().max()) == 0.0
    z = torch.as_tensor(_potentials(L, 31, L) * 1e-11, device="cuda")
    y = st.coeffs.clone()
    y[:8] += z
    out = m.tendency(y)
    bve = BarotropicVorticity(pl, viscosity=0.0).tendency(BarotropicState(z))
    scale = float(bve.abs().max())
    err = max(float((out[k] - bve).abs().max()) for k in range(8)) / scale
    print(f"  bve degeneracy {err:.2e}", flush=True)
    assert err <= 1e-12


def test_vector_curl_div_closed_forms_at_L85():
    """Gauss 128 x 256 (3/2-rule product), L = 85, random potentials in every
    mode up to L - 1: curl / div of k x grad(psi) + grad(chi) recover
    lap(psi), lap(chi) within 5e-12 (the formulation's own round-off at this
    size: the oracle's closed-form errors for these potentials are 6.6e-13 /
    1.8e-12) and equal the oracle's within 1e-12."""
    import torch
    from aeolus_amd.planet import Planet
    from oracle.models import planet_constants
    R, om = planet_constants(24.0)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        pl = Planet.with_radius(R, om, grid_type="latlon", nlat=128, nlon=256, l_max=85,
                                product_quadrature="fine")
    so = pl.so
    sh_p = so._product_space.sh
    psi = torch.as_tensor(_potentials(85, 41, 84), device="cuda")
    chi = torch.as_tensor(_potentials(85, 42, 84), device="cuda")
    c = so._product_space.coslat

    def derivs(a):
        return (sh_p.inv_transform(so.d_lambda_coeffs(a)),
                sh_p.inv_transform(so.sin_theta_d_theta_coeffs(a)) / so.R)
    p_l, p_s = derivs(psi)
    c_l, c_s = derivs(chi)
    fe, fn = (p_s + c_l) / c, (p_l - c_s) / c
    curl, div = so.vector_curl_div_spectral(fe, fn, truncate=False)
    lap = so.lap_eigs[:, None]
    scale = max(float((lap * psi).abs().max()), float((lap * chi).abs().max()))
    ec = float((curl - lap * psi).abs().max()) / scale
    ed = float((div - lap * chi).abs().max()) / scale
    print(f"case vector curl/div L=85 closed forms: curl {ec:.2e} div {ed:.2e}", flush=True)
    assert ec < 5e-12 and ed < 5e-12
    from oracle import pe
    from oracle.models import Planet as OPlanet
    op = OPlanet(l_max=85, radius=R, omega=om, grid_type="latlon", nlat=128, nlon=256)
    rc, rd = pe.vector_curl_div(op, pe.extended_transform(op), _np(fe), _np(fn), truncate=False)
    assert np.abs(_np(curl) - rc).max() < 1e-12 * scale
    assert np.abs(_np(div) - rd).max() < 1e-12 * scale


def test_vector_curl_div_coarse_quadrature_matches_oracle():
    """The coarse product quadrature (the state sampling) builds the
    extended plan on the state grid; against the oracle within 1e-12."""
    import torch
    from aeolus_amd.planet import Planet
    from oracle import pe
    from oracle.models import Planet as OPlanet, planet_constants
    R, om = planet_constants(24.0)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        pl = Planet.with_radius(R, om, grid_type="latlon", nlat=32, nlon=64, l_max=15,
                                product_quadrature="coarse")
    op = OPlanet(l_max=15, radius=R, omega=om, grid_type="latlon", nlat=32, nlon=64,
                 product_quadrature="coarse")
    psi, chi = _potentials(15, 3, 14), _potentials(15, 4, 14)
    fe, fn = _vector(op, psi, chi)
    rc, rd = pe.vector_curl_div(op, pe.extended_transform(op), fe, fn)
    gc, gd = pl.so.vector_curl_div_spectral(torch.as_tensor(fe, device="cuda"),
                                            torch.as_tensor(fn, device="cuda"))
    scale = max(np.abs(rc).max(), np.abs(rd).max())
    assert np.abs(_np(gc) - rc).max() < 1e-12 * scale
    assert np.abs(_np(gd) - rd).max() < 1e-12 * scale


def test_vector_curl_div_batch_larger_than_plan_chunk():
    """A 20-field batch (40 analyses, more than the extended plan's 16 fields
    per launch sequence) equals field-by-field calls."""
    import torch
    pl, op = _planets("latlon")
    N = pl.so.extended_product_transform().n_points
    rng = np.random.default_rng(77)
    fe = torch.as_tensor(rng.standard_normal((20, N)), device="cuda")
    fn = torch.as_tensor(rng.standard_normal((20, N)), device="cuda")
    bc, bd = pl.so.vector_curl_div_spectral(fe, fn)
    for f in (0, 7, 15, 16, 19):
        sc, sd = pl.so.vector_curl_div_spectral(fe[f], fn[f])
        assert torch.allclose(bc[f], sc, rtol=0, atol=1e-14 * float(sc.abs().max()))
        assert torch.allclose(bd[f], sd, rtol=0, atol=1e-14 * float(sd.abs().max()))
