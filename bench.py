"""Benchmark: spectral RK4 steps/s on one MI355X.

Headline (``value``): barotropic-vorticity RK4 steps/s at L = 170 on the
Gauss 256 x 512 grid (single member, dt = 300 s, seeded l^-1.5 vorticity
with rms 5e-5 s^-1, R = 6.371e6 m, 24 h rotation) for the fastest precision
mode whose vorticity after one simulated day (288 steps) stays within 1e-11
relative L2 of the plain FP64 step: plain FP64 Legendre contractions, or
split INT8-digit MFMA contractions with k in {4, 5, 6, 7} slices (FP64 state,
recombination, tables and RK4 update).  The plain FP64 step itself matches
the NumPy oracle to <= 1e-12 (tests).

Measured in the same run (``secondary``):
  * the same at L = 85 on Gauss 128 x 256;
  * shallow water at L = 85 on Gauss 128 x 256 with a 2 km Gaussian mountain,
    FP64 and split, per-field error after one day;
  * the uneven-L ensemble: 64 members each at L = 21/42/85/170 on
    32x64/64x128/128x256/256x512, one hipGraph replay per two steps,
    autotuned tiles, FP64 and split, with the one-day error;
  * the adjoint: L = 170, 48 steps, split transforms (7 slices); cost of the
    checkpointed forward + reverse sweep relative to the plain forward;
  * the geodesic res 4 / L = 21 trajectory with FP64 and split (6 slices)
    dense Y-basis contractions.

Inputs are resident in HBM before every timed region; a "step" is one full
RK4 step (4 fused stages).  Prints one JSON line.  ``--gpus N`` with N > 1
exits: every figure is one GPU's.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time
import warnings

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

R_EARTH = 6.371e6
OMEGA = 2.0 * math.pi / 86400.0
HBM_PEAK_GBS = 8000.0
DAY_STEPS = 288          # one simulated day at dt = 300 s
DT = 300.0
TOL = 1e-11
SPLITS = (4, 5, 6, 7)
def kclass(cls: int, nslices: int) -> str:
    """The kernel a single-member BVE step launches for a profiling class
    (0 synthesis prime, 1 grid stage, 2 Legendre stage) at a precision."""
    if cls == 0:
        return ("bve1_stage_kernel (synthesis only)" if nslices == 0
                else "leg_stage_kernel<KindBve, EpiNoneS> (synthesis only)")
    if cls == 1:
        return ("grid_jac_hemi_kernel (per-row FFTs + Jacobian)" if nslices == 0
                else "grid_jac_p2_kernel (row-pair FFTs + Jacobian)")
    return ("bve1_stage_kernel (analysis + RK4 epilogue + next synthesis)" if nslices == 0
            else "leg_stage_kernel<KindBve, EpiBveS> (analysis + RK4 epilogue + next synthesis)")


def _args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=576)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--headline-only", action="store_true",
                   help="only the headline workload (for profiler runs)")
    p.add_argument("--precision", type=int, default=None,
                   help="force the headline precision mode (0 = FP64, k = split slices)")
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run 4 device copies of 256 MiB (scales profiler byte counters)")
    p.add_argument("--pe-only", nargs="?", const="all", default=None,
                   choices=["all", "geodesic", "gauss"],
                   help="only the primitive-equation tendency / RK4 timings (one JSON line)")
    p.add_argument("--ensemble-only", action="store_true",
                   help="only the 256-member uneven-L ensemble, graph replay (for profiler "
                        "runs; precision from --precision, default 6 slices)")
    p.add_argument("--batch-only", type=int, nargs="?", const=170, default=None, metavar="L",
                   help="one 64-member BVE batch at L = 170 (or 85) alone: `--steps` RK4 "
                        "steps after `--warmup`, `--precision` and the first `--tiles` label "
                        "(default: autotune), one JSON line (profiler runs)")
    p.add_argument("--tiles", type=str, default=None,
                   help="tiles of the L = 21,42,85,170 classes, fused members per workgroup "
                        "or the GEMM form, e.g. mb8,mb8,gemm64x8x2,gemm64x8x2 (default: autotune)")
    return p.parse_args()


def copy_calibration():
    """Plain copy kernel of known size: FETCH_SIZE / WRITE_SIZE per dispatch of
    this kernel against 256 MiB gives the counters' scale."""
    import torch
    n = 256 * 1024 * 1024 // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(4):
        b.copy_(a)
    _sync()
    print(json.dumps({"copy_bytes_per_dispatch": n * 8}))


def _planet(L, nlat, nlon):
    from aeolus_amd.planet import Planet
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return Planet.with_radius(R_EARTH, OMEGA, grid_type="latlon", nlat=nlat, nlon=nlon,
                                  l_max=L, product_quadrature="fine")


def _rel(a, b):
    import torch
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def _sync():
    import torch
    torch.cuda.synchronize()


def _timed(run, steps, warmup):
    """ms per step of run(k) (k RK4 steps), events on the library's stream."""
    import torch
    if warmup:
        run(warmup)
    _sync()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    run(steps)
    t1.record()
    _sync()
    return t0.elapsed_time(t1) / steps


def _profiled(model, cls, run, steps):
    """(ms per step, mean launch ms, launches) of kernel class `cls` over `steps`."""
    from aeolus_amd import _native as nat
    lib = nat.lib()
    nat.check(lib.aeo_model_profile(model._model, cls, 4 * steps + 8), "profile")
    try:
        ms = _timed(run, steps, 0)
        tot, cnt = ctypes.c_double(), ctypes.c_int()
        nat.check(lib.aeo_model_profile_read(model._model, ctypes.byref(tot), ctypes.byref(cnt)),
                  "profile_read")
    finally:
        nat.check(lib.aeo_model_profile(model._model, -1, 0), "profile off")
    return ms, tot.value / max(cnt.value, 1), cnt.value


def _split_cols(L, cus=256):
    """Columns the single-member FP64 stage splits into two degree halves
    (bve1_msplit: the CUs left over by one workgroup per column, each half
    keeping at least 8 degrees)."""
    n = L + 1
    return max(0, min(cus - n, n - 16))


def _stage_bytes(cls, L, nlat, nslices):
    """Algorithmic HBM bytes of one launch of a BVE stage kernel class:
    tables and intermediates read or written once, state passes.  FP64
    synthesis reads no table rows beyond the recurrence seeds (2 rows per
    degree block: 4 blocks per column, 8 in a column split into two degree
    halves) and the (a, b) recurrence and C+/C- columns; split-precision
    synthesis reads its digit table (nslices INT8 digits per entry).  In FP64
    the split columns' E/O arrive in two planes (written by the stage, read
    by the grid stage) and the grid stage writes the unweighted row spectra
    FN/FS in place of S/D (the same bytes)."""
    nh = (nlat + 1) // 2
    n = L + 1
    B = (L + 1) * (L + 2) // 2
    cut = (2 * L) // 3
    Bc = (cut + 1) * (cut + 2) // 2
    spec = n * n * 16
    tab = 8 if nslices == 0 else nslices          # bytes per table entry
    ms = _split_cols(L) if nslices == 0 else 0
    eo = 2 * 4 * nh * (n + ms) * 16               # E/O, 4 fields (+ plane 1 of split columns)
    sd = 2 * 1 * nh * (cut + 1) * 16              # S/D (or FN/FS), 1 field, m <= cut
    cols = n * n * (16 + 8 + 8)                   # (a, b), C+, C- columns
    syn = nh * B * tab if nslices else 4 * 2 * nh * (n + ms) * 8
    if cls == 0:
        return syn + cols + spec + eo
    if cls == 1:
        return eo + sd
    # analysis table (truncated) + S/D + ys, y0, ksum r/w, ynext + synthesis + E/O
    return nh * Bc * tab + sd + 5 * spec + syn + cols + eo


def _stage_flops(cls, L, nlat):
    """Algorithmic flops of one launch (2 nlat B per Legendre field transform)."""
    B = (L + 1) * (L + 2) // 2
    cut = (2 * L) // 3
    Bc = (cut + 1) * (cut + 2) // 2
    if cls == 0:
        return 4 * 2 * nlat * B
    if cls == 1:
        n = 2 * (L + 1)
        return 5 * 2.5 * n * math.log2(max(n, 2)) * nlat
    return 2 * nlat * Bc + 4 * 2 * nlat * B


def bve_modes(L, nlat, nlon, steps, warmup, seed, force=None):
    """Per precision mode: ms/step over `steps` and the one-day error vs FP64."""
    import torch
    from aeolus_amd.physics import BarotropicVorticity
    from aeolus_amd.run.initial_conditions import spectrum_l15
    pl = _planet(L, nlat, nlon)
    model = BarotropicVorticity(pl)
    z0 = torch.as_tensor(spectrum_l15(L, np.random.default_rng(seed), rms=5e-5), device="cuda")
    out = torch.empty_like(z0)
    run = lambda k: model.rk4(z0, DT, k, out=out)
    modes = {}
    ref_day = model.rk4(z0, DT, DAY_STEPS)
    for ns in (0,) + SPLITS:
        if force is not None and ns not in (0, force):
            continue
        model.set_precision(ns)
        day = model.rk4(z0, DT, DAY_STEPS)
        err = 0.0 if ns == 0 else _rel(day, ref_day)
        ms = _timed(run, min(steps, 192), warmup)
        modes[ns] = {"ms_per_step": ms, "rel_l2_vs_fp64_1day": err,
                     "finite": bool(torch.isfinite(day).all())}
    ok = [ns for ns, v in modes.items() if v["rel_l2_vs_fp64_1day"] <= TOL and v["finite"]]
    best = force if force is not None else min(ok, key=lambda ns: modes[ns]["ms_per_step"])
    return model, run, modes, best


def headline(args):
    model, run, modes, best = bve_modes(170, 256, 512, args.steps, args.warmup, 2026,
                                        args.precision)
    model.set_precision(best)
    cls_ms = {}
    for c in (1, 2):
        _, kms, _ = _profiled(model, c, run, 24)
        cls_ms[c] = kms
    dom = max(cls_ms, key=cls_ms.get)
    # the timed region: W warm-up steps, then exactly K steps (uninstrumented)
    ms_step = _timed(run, args.steps, args.warmup)
    # the same K steps again with the dominant kernel's launches timed by
    # their own dispatch events (hipExtLaunchKernelGGL); the event packets
    # serialise the launches, so this pass is slower and only gives durations;
    # three passes, the one with the least mean launch duration kept (the
    # event-timed durations vary between runs by up to 15 %)
    ms_instr, kms, nl = min((_profiled(model, dom, run, args.steps) for _ in range(3)),
                            key=lambda r: r[1])
    return dict(model=model, modes=modes, best=best, dom=dom, ms_step=ms_step,
                ms_instr=ms_instr, dom_ms=kms, dom_launches=nl, cls_ms=cls_ms)


def swe_modes(steps, warmup):
    import torch
    from aeolus_amd.physics import ShallowWaterModel, Topography
    from aeolus_amd.run.initial_conditions import swe_spectra
    L, nlat, nlon = 85, 128, 256
    pl = _planet(L, nlat, nlon)
    rng = np.random.default_rng(85)
    lat0 = math.degrees(math.asin(rng.uniform(-1, 1)))
    lon0 = math.degrees(rng.uniform(0, 2 * math.pi))
    topo = Topography.mountain(pl, height_m=2000.0, lat_deg=lat0, lon_deg=lon0,
                               width_deg=math.degrees(1e6 / R_EARTH))
    swe = ShallowWaterModel(pl, gravity=9.80616, mean_depth=1e4, topography=topo)
    y0 = torch.as_tensor(swe_spectra(L, R_EARTH, 9.80616, seed=85), device="cuda")
    out = torch.empty_like(y0)
    run = lambda k: swe.rk4(y0, DT, k, out=out)
    ref = swe.rk4(y0, DT, DAY_STEPS)
    res = {}
    for ns in (0,) + SPLITS:
        swe.set_precision(ns)
        day = swe.rk4(y0, DT, DAY_STEPS)
        errs = [0.0 if ns == 0 else _rel(day[f], ref[f]) for f in range(3)]
        res[ns] = {"ms_per_step": _timed(run, min(steps, 192), warmup),
                   "rel_l2_vs_fp64_1day": dict(zip(("zeta", "delta", "phi"), errs)),
                   "finite": bool(torch.isfinite(day).all())}
    return res


def ensemble_modes(steps, ns_split):
    import torch
    from aeolus_amd.run.ensemble import Ensemble
    from aeolus_amd.run.initial_conditions import spectrum_l15
    classes = [(21, 32, 64), (42, 64, 128), (85, 128, 256), (170, 256, 512)]
    nper = 64
    planets = [_planet(L, a, b) for L, a, b in classes]
    states = [torch.as_tensor(np.stack([spectrum_l15(L, np.random.default_rng(1000 * L + k),
                                                     rms=5e-5) for k in range(nper)]),
                              device="cuda") for L, _, _ in classes]
    res, ref = {}, None
    for ns in (0, ns_split):
        ens = Ensemble([(pl, nper) for pl in planets], precision=ns)
        tuning = ens.autotune(states, DT, steps=2)
        outs = [torch.empty_like(s) for s in states]
        run = lambda k: ens.rk4(states, DT, k, out=outs, use_graph=True)
        ms = _timed(run, steps, 4)
        day = ens.rk4(states, DT, DAY_STEPS)
        if ns == 0:
            ref = day
        errs = {f"L{L}": (0.0 if ns == 0 else _rel(d, r))
                for (L, _, _), d, r in zip(classes, day, ref)}
        res[ns] = {"ms_per_ensemble_step": ms, "member_steps_per_s": 4 * nper * 1000.0 / ms,
                   "rel_l2_vs_fp64_1day": errs,
                   "tiles": {f"L{L}": _tile_label(t) for (L, _, _), t in zip(classes, tuning)}}
        del ens
    return res


def _tile_label(t) -> str:
    """Autotuner choice as a label: mb<k> (fused, k members per workgroup) or
    gemm<cols>x<members>x<tiles> (GEMM form)."""
    return f"mb{t['mb']}" if t["mb"] else "gemm" + "x".join(str(v) for v in t["gemm"])


def _apply_tile(model, label: str) -> None:
    if label.startswith("gemm"):
        model.set_tiles(0)
        model.set_gemm_tiles(*(int(v) for v in label[4:].split("x")))
    else:
        model.set_tiles(int(label[2:] if label.startswith("mb") else label))


def ensemble_only(args):
    """The ensemble workload alone: `--steps` graph-replayed RK4 steps of the
    256 members after `--warmup`, one JSON line (profiler runs)."""
    import torch
    from aeolus_amd.run.ensemble import Ensemble
    from aeolus_amd.run.initial_conditions import spectrum_l15
    classes = [(21, 32, 64), (42, 64, 128), (85, 128, 256), (170, 256, 512)]
    nper = 64
    ns = 6 if args.precision is None else args.precision
    planets = [_planet(L, a, b) for L, a, b in classes]
    states = [torch.as_tensor(np.stack([spectrum_l15(L, np.random.default_rng(1000 * L + k),
                                                     rms=5e-5) for k in range(nper)]),
                              device="cuda") for L, _, _ in classes]
    ens = Ensemble([(pl, nper) for pl in planets], precision=ns)
    if args.tiles:
        tiles = args.tiles.split(",")
        for m, t in zip(ens.models, tiles):
            _apply_tile(m, t)
    else:
        tiles = [_tile_label(t) for t in ens.autotune(states, DT, steps=2)]
    outs = [torch.empty_like(s) for s in states]
    ms = _timed(lambda k: ens.rk4(states, DT, k, out=outs, use_graph=True), args.steps,
                args.warmup)
    print(json.dumps({"workload": "BVE ensemble, 64 members each at L = 21/42/85/170",
                      "precision": _mode_name(ns), "tiles": tiles,
                      "steps": args.steps, "ms_per_ensemble_step": ms,
                      "member_steps_per_s": 4 * nper * 1000.0 / ms}))


def batch_only(args):
    """One 64-member BVE batch (the L = 170 or L = 85 class of the ensemble) on
    its own: `--steps` fused RK4 steps after `--warmup`, one JSON line."""
    import torch
    from aeolus_amd.physics.barotropic import BarotropicBatch
    from aeolus_amd.autotune import tune_members_per_block
    from aeolus_amd.run.initial_conditions import spectrum_l15
    L = int(args.batch_only)
    grids = {170: (256, 512), 85: (128, 256)}
    if L not in grids:
        raise SystemExit("--batch-only: L must be 170 or 85")
    nper = 64
    ns = 6 if args.precision is None else args.precision
    model = BarotropicBatch(_planet(L, *grids[L]), nper)
    model.set_precision(ns)
    y = torch.as_tensor(np.stack([spectrum_l15(L, np.random.default_rng(1000 * L + k),
                                               rms=5e-5) for k in range(nper)]), device="cuda")
    if args.tiles:
        label = args.tiles.split(",")[0]
        _apply_tile(model, label)
    else:
        label = _tile_label(tune_members_per_block(model, y, DT, steps=2))
    out = torch.empty_like(y)
    ms = _timed(lambda k: model.rk4(y, DT, k, out=out), args.steps, args.warmup)
    print(json.dumps({"workload": f"BVE batch, 64 members at L = {L}",
                      "precision": _mode_name(ns), "tiles": label, "steps": args.steps,
                      "ms_per_step": ms, "member_steps_per_s": nper * 1000.0 / ms}))


def geodesic_modes(steps, warmup):
    """Geodesic res 4 state / res 5 product grid, L = 21, 96 x 900 s (the C2
    trajectory): ms/step and the error against FP64 for the dense Y-basis
    contractions in FP64 and split precision (6 slices)."""
    import torch
    from aeolus_amd.physics import BarotropicVorticity
    from aeolus_amd.planet import Planet
    from aeolus_amd.run.initial_conditions import spectrum_l15
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        pl = Planet.with_radius(R_EARTH, OMEGA, grid_type="geodesic", grid_resolution=4,
                                l_max=21, product_quadrature="fine")
    model = BarotropicVorticity(pl)
    z0 = torch.as_tensor(spectrum_l15(21, np.random.default_rng(21), rms=5e-5), device="cuda")
    out = torch.empty_like(z0)
    res, ref = {}, None
    for ns in (0, 6):
        model.set_precision(ns)
        zT = model.rk4(z0, 900.0, 96)
        ref = zT if ns == 0 else ref
        ms = _timed(lambda k: model.rk4(z0, 900.0, k, out=out), min(steps, 96), warmup)
        res[_mode_name(ns)] = {"ms_per_step": ms,
                               "rel_l2_vs_fp64_96steps": 0.0 if ns == 0 else _rel(zT, ref)}
    return res


def _pe_state(L, K, seed):
    """Isothermal rest (260 K, 101325 Pa) plus seeded perturbations band-limited
    at the 2/3 cut (zeta, delta 2e-6 1/s, T 0.5 K, ln p_s 1e-3)."""
    n, cut = L + 1, (2 * L) // 3
    rng = np.random.default_rng(seed)
    y = np.zeros((3 * K + 1, n, n), dtype=np.complex128)
    y[2 * K:3 * K, 0, 0] = 260.0 * math.sqrt(4.0 * math.pi)
    y[3 * K, 0, 0] = math.log(101325.0) * math.sqrt(4.0 * math.pi)

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


def pe_modes(steps, which="all"):
    """Dry primitive-equation tendency (8 uniform sigma levels) and one RK4
    step through rk4_step_array: on the reference's default PE configuration
    (geodesic res 3 state / res 4 product, L = 10) and on Gauss 128 x 256 /
    L = 85 and 256 x 512 / L = 170 grids (3/2-rule products)."""
    import torch
    from aeolus_amd.physics import PrimitiveEquationsModel, SigmaGrid
    from aeolus_amd.planet import Planet
    from aeolus_amd.run.engine import rk4_step_array
    res = {}
    for name, kw, L in (("geodesic_res3_L10_K8", dict(grid_type="geodesic", grid_resolution=3), 10),
                        ("gauss_128x256_L85_K8", dict(grid_type="latlon", nlat=128, nlon=256), 85),
                        ("gauss_256x512_L170_K8", dict(grid_type="latlon", nlat=256, nlon=512), 170)):
        if which != "all" and not name.startswith(which):
            continue
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            pl = Planet.with_radius(R_EARTH, OMEGA, l_max=L, product_quadrature="fine", **kw)
        m = PrimitiveEquationsModel(pl, SigmaGrid.uniform(8))
        y = torch.as_tensor(_pe_state(L, 8, 7), device="cuda")
        ms_t = _timed(lambda k: [m.tendency(y) for _ in range(k)], steps, 5)
        ms_rk = _timed(lambda k: [rk4_step_array(m.tendency, y, 0.0, 300.0) for _ in range(k)],
                       max(steps // 4, 1), 2)
        res[name] = {"ms_per_tendency": ms_t, "ms_per_rk4_step": ms_rk,
                     "product_points": m._ext.n_points, "levels": 8,
                     "fields_per_tendency": {"synthesised": 82, "analysed_L": 33,
                                             "analysed_L_plus_1": 16}}
    return res


def adjoint_cost(model, ns):
    import torch
    from aeolus_amd.run.initial_conditions import spectrum_l15
    model.set_precision(ns)
    n = 48
    z0 = torch.as_tensor(spectrum_l15(170, np.random.default_rng(5), rms=5e-5), device="cuda")
    yb = torch.as_tensor(spectrum_l15(170, np.random.default_rng(6), rms=5e-5), device="cuda")
    out = torch.empty_like(z0)
    reps = 5
    fwd = _timed(lambda k: [model.rk4(z0, DT, n, out=out) for _ in range(k)], reps, 1)
    _, ck = model.rk4_record(z0, DT, n)
    rec = _timed(lambda k: [model.rk4_record(z0, DT, n, out=out) for _ in range(k)], reps, 1)
    adj = _timed(lambda k: [model.rk4_adjoint(ck, DT, n, yb, out=out) for _ in range(k)], reps, 1)
    return {"steps": n, "precision_slices": ns, "forward_ms": fwd, "record_ms": rec,
            "adjoint_ms": adj, "adjoint_to_forward": adj / fwd,
            "gradient_to_forward": (rec + adj) / fwd,
            "checkpoint_MB": model.checkpoint_elems(n) * 16 / 1e6}


def cpu_baseline(seconds: float):
    """Oracle restatement (NumPy FFT + stable Legendre) of the same L = 170 BVE step."""
    from oracle.models import Planet as OPlanet, bve_tendency, rk4_step
    from aeolus_amd.run.initial_conditions import spectrum_l15
    blas = _blas_threads()
    op = OPlanet(l_max=170, radius=R_EARTH, omega=OMEGA, grid_type="latlon", nlat=256,
                 nlon=512)
    y = spectrum_l15(170, np.random.default_rng(2026), rms=5e-5)
    tend = lambda s: bve_tendency(op, s)
    y = rk4_step(tend, y, DT)  # warm-up step
    n = 0
    t0 = time.perf_counter()
    while True:
        y = rk4_step(tend, y, DT)
        n += 1
        if time.perf_counter() - t0 >= seconds:
            break
    el = time.perf_counter() - t0
    pe_ms = _pe_oracle_ms(min(3.0, seconds / 4))
    return {"value": n / el, "unit": "RK4 steps/s", "cores": int(blas),
            "pe_geodesic_res3_L10_K8_ms_per_tendency": pe_ms,
            "host_cores_available": len(os.sched_getaffinity(0)), "kind": "port",
            "sample": f"{n} RK4 steps of the L=170 Gauss 256x512 BVE step (oracle NumPy "
                      f"FFT + stable-Legendre restatement, FP64), after 1 warm-up step"}


def _blas_threads() -> int:
    try:
        from threadpoolctl import threadpool_info
        return int(max((i.get("num_threads", 1) for i in threadpool_info()), default=1))
    except Exception:
        return 1


def reference_algorithm_cpu(c3_steps: int = 2, progress=lambda msg: None):
    """The reference algorithm itself on the host cores: dense basis Y built
    with the reference kernel's formula (cu/sh_matrix.cu:7-76), analysis
    Y^H (w f) and synthesis 2 Re(Y a) - Re(Y a_0) per field
    (num/fast_geodesic_sh.py:68-107), called the way the tendencies call them
    (phy/barotropic.py:120-167, phy/shallow_water.py:347-418), RK4 of
    run/bve/runner.py:211-216 / run/engine.py:406-431 -- the oracle's NumPy
    restatement with basis="reference", on the same seeded inputs as the
    device runs:
      C1  BVE Gauss 12x24 / 13x25 product, L = 8, 144 x 600 s (all steps);
      C2  BVE geodesic res 4 / res 5 product, L = 21, 96 x 900 s (all steps);
      C3  SWE Gauss 128x256, L = 85, 2 km mountain, c3_steps x 300 s (dense Y
          is 1.96 GB per grid; the mountain's analysis and synthesis run first
          and warm both grids' Y).
    C1 and C2 are timed after one warm-up step."""
    from aeolus_amd.run.initial_conditions import gaussian_vortex_pair, spectrum_l15, swe_spectra
    from oracle.models import (Planet as OPlanet, SWEParams, bve_tendency, mountain, rk4_step,
                               rk4_step_array, swe_tendency)
    res = {"algorithm": "dense Y (cu/sh_matrix.cu formula), zgemv-shaped NumPy transforms",
           "blas_threads": _blas_threads(), "host_cores_available": len(os.sched_getaffinity(0))}

    def run_bve(name, p, z, dt, n):
        tend = lambda y: bve_tendency(p, y)
        z = rk4_step(tend, z, dt)                                   # warm-up step
        t0 = time.perf_counter()
        for _ in range(n):
            z = rk4_step(tend, z, dt)
        el = time.perf_counter() - t0
        res[name] = {"steps": n, "seconds": el, "steps_per_s": n / el}

    p = OPlanet(l_max=8, radius=R_EARTH, omega=OMEGA, grid_type="latlon", nlat=12, nlon=24,
                dense=True, basis="reference")
    z = p.sh.transform(gaussian_vortex_pair(p.grid.point_latitudes, p.grid.point_longitudes,
                                            R_EARTH, seed=1))
    run_bve("c1_bve_gauss12x24_L8", p, z, 600.0, 144)
    progress("CPU reference algorithm C1 done")
    p = OPlanet(l_max=21, radius=R_EARTH, omega=OMEGA, grid_type="geodesic", resolution=4,
                basis="reference")
    run_bve("c2_bve_geodesic_res4_L21", p, spectrum_l15(21, np.random.default_rng(21), rms=5e-5),
            900.0, 96)
    progress("CPU reference algorithm C2 done")
    t0 = time.perf_counter()
    p = OPlanet(l_max=85, radius=R_EARTH, omega=OMEGA, grid_type="latlon", nlat=128, nlon=256,
                dense=True, basis="reference")
    build = time.perf_counter() - t0
    rng = np.random.default_rng(85)
    lat0 = math.degrees(math.asin(rng.uniform(-1, 1)))
    lon0 = math.degrees(rng.uniform(0, 2 * math.pi))
    elev = mountain(p, height_m=2000.0, lat_deg=lat0, lon_deg=lon0,
                    width_deg=math.degrees(1e6 / R_EARTH))
    prm = SWEParams(p, gravity=9.80616, mean_depth=1e4, elevation_lm=elev)
    y = swe_spectra(85, R_EARTH, 9.80616, seed=85)
    progress("CPU reference algorithm C3 basis built")
    t0 = time.perf_counter()
    for k in range(c3_steps):
        y = rk4_step_array(lambda s: swe_tendency(prm, s), y, DT)
        progress(f"CPU reference algorithm C3 step {k + 1}")
    el = time.perf_counter() - t0
    res["c3_swe_gauss128x256_L85_mountain"] = {"steps": c3_steps, "seconds": el,
                                               "steps_per_s": c3_steps / el,
                                               "basis_build_s": build}
    return res


def c1_device(steps: int):
    """C1 on the device: BVE Gauss 12x24 / 13x25 product, L = 8, FP64, the
    two-vortex start of the reference algorithm's CPU run; ms per RK4 step."""
    import torch
    from aeolus_amd.physics import BarotropicVorticity
    from aeolus_amd.planet import Planet
    from aeolus_amd.run.initial_conditions import gaussian_vortex_pair
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        pl = Planet.with_radius(R_EARTH, OMEGA, grid_type="latlon", nlat=12, nlon=24, l_max=8,
                                product_quadrature="fine")
    model = BarotropicVorticity(pl)
    lat = pl.grid.point_latitudes
    lon = pl.grid.point_longitudes
    lat = lat.cpu().numpy() if hasattr(lat, "cpu") else np.asarray(lat)
    lon = lon.cpu().numpy() if hasattr(lon, "cpu") else np.asarray(lon)
    z0 = pl.sh.transform(torch.as_tensor(gaussian_vortex_pair(lat, lon, R_EARTH, seed=1),
                                         device="cuda"))
    out = torch.empty_like(z0)
    return _timed(lambda k: model.rk4(z0, 600.0, k, out=out), steps, 16)


def _pe_oracle_ms(seconds: float) -> float:
    """Oracle primitive-equation tendency (NumPy, dense stable-recurrence basis)
    on the reference's default PE case: geodesic res 3 (res 4 product), L = 10,
    8 uniform levels; ms per tendency over about `seconds`."""
    from oracle import pe
    from oracle.models import Planet as OPlanet
    op = OPlanet(l_max=10, radius=R_EARTH, omega=OMEGA, grid_type="geodesic", resolution=3)
    m = pe.PEModel(op, pe.Sigma.uniform(8))
    y = _pe_state(10, 8, 7)
    pe.pe_tendency(m, y)
    n, t0 = 0, time.perf_counter()
    while n == 0 or time.perf_counter() - t0 < seconds:
        pe.pe_tendency(m, y)
        n += 1
    return (time.perf_counter() - t0) * 1000.0 / n


def kernel_sources_hash() -> str:
    """Short SHA-256 of every HIP source of the library (csrc/*.hip, *.hpp):
    a traffic profile records it, and a profile taken before the kernels
    changed is not used."""
    import glob
    import hashlib
    h = hashlib.sha256()
    for path in sorted(glob.glob(os.path.join(ROOT, "aeolus_amd", "csrc", "*.*p*"))):
        if path.endswith((".hip", ".hpp")):
            h.update(os.path.basename(path).encode())
            with open(path, "rb") as f:
                h.update(f.read())
    return h.hexdigest()[:16]


def _measured_traffic(kernel: str, precision: str):
    """(bytes, status): HBM bytes per launch of `kernel` from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of the same workload
    (profiles/*traffic*.md, counters scaled by a plain copy of known size),
    provided the summary was taken from the current kernel sources; else
    (None, why)."""
    import glob
    import re
    want = kernel_sources_hash()
    stale = False
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*traffic*.md"))):
        with open(path) as f:
            for line in f:
                m = re.match(r"traffic_bytes_per_launch: (\d+) kernel: (.+?) precision: (\S+)"
                             r"(?: workload: \S+)?(?: sources: (\S+))?$", line.strip())
                if m and kernel.startswith(m.group(2)) and m.group(3) == precision:
                    if m.group(4) == want:
                        return int(m.group(1)), "current"
                    stale = True
    return None, ("stale: the traffic profile predates the current kernel sources" if stale
                  else "no traffic profile for this kernel and precision")


def _mode_name(ns):
    return "fp64" if ns == 0 else f"split{ns}"


def _progress(msg: str) -> None:
    """One progress line per finished part on stderr (stdout keeps the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = _args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.ensemble_only:
        ensemble_only(args)
        return
    if args.batch_only:
        batch_only(args)
        return
    if args.pe_only:
        print(json.dumps({"workload": "dry primitive equations, 8 sigma levels",
                          "pe_modes": pe_modes(min(args.steps, 100), args.pe_only)}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "every figure of this benchmark is one GPU's; "
                                   "run with --gpus 1"}))
        sys.exit(2)
    h = headline(args)
    _progress("headline L=170 done")
    best, dom = h["best"], h["dom"]
    nbytes = _stage_bytes(dom, 170, 256, best)
    achieved = nbytes / (h["dom_ms"] * 1e-3) / 1e9
    flops = _stage_flops(dom, 170, 256)
    traffic = _measured_traffic(kclass(dom, best), _mode_name(best))
    secondary = {
        "bve170_modes": {_mode_name(k): v for k, v in h["modes"].items()},
        "bve170_ms_per_step_kernel_timed_pass": h["ms_instr"],
        "bve170_kernel_class_ms": {kclass(k, best): v for k, v in h["cls_ms"].items()},
        "dominant_kernel_gflops_per_s": flops / (h["dom_ms"] * 1e-3) / 1e9,
    }
    if not args.headline_only:
        _, _, m85, b85 = bve_modes(85, 128, 256, args.steps, args.warmup, 85)
        _progress("L=85 modes done")
        fp = h["modes"][0]["ms_per_step"]
        secondary["bve170_split_speedup_vs_fp64"] = {
            _mode_name(k): fp / v["ms_per_step"] for k, v in h["modes"].items() if k}
        secondary["bve85_modes"] = {_mode_name(k): v for k, v in m85.items()}
        secondary["bve85_best"] = _mode_name(b85)
        secondary["swe85_mountain_modes"] = {_mode_name(k): v
                                             for k, v in swe_modes(args.steps, 5).items()}
        ns_split = best if best else 6
        secondary["ensemble_256_members"] = {_mode_name(k): v for k, v in
                                             ensemble_modes(48, ns_split).items()}
        _progress("SWE and ensemble done")
        secondary["adjoint_L170"] = adjoint_cost(h["model"], 7)
        # the same with FP64 contractions (the forward on the single-member kernels)
        secondary["adjoint_L170_fp64"] = adjoint_cost(h["model"], 0)
        secondary["geodesic_res4_L21_modes"] = geodesic_modes(args.steps, args.warmup)
        secondary["primitive_equations_K8"] = pe_modes(min(args.steps, 100))
        secondary["c1_bve_gauss12x24_L8_fp64_ms_per_step"] = c1_device(144)
        _progress("device secondaries done")
    line = {
        "metric": "RK4 steps/s at lmax 85/170 vs plain FP64 at rel-err<=1e-11; adjoint:fwd cost",
        "value": 1000.0 / h["ms_step"],
        "unit": "RK4 steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["ms_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if best == 0 else f"int8 digits x{best} (int32 accum) + f64",
        "data": "synthetic (seeded l^-1.5 spectra, random-phase)",
        "config": {"workload": f"BVE RK4, Gauss 256x512, L=170, single member, dt=300 s, "
                               f"{_mode_name(best)} Legendre contractions",
                   "grid": "256x512", "l_max": 170, "dt_s": DT, "radius_m": R_EARTH,
                   "precision": _mode_name(best), "accuracy_bar": "rel-L2 <= 1e-11 vs FP64 "
                   "after 1 simulated day"},
        "secondary": secondary,
        "roofline": {"bound": "hbm", "kernel": kclass(dom, best), "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic[0], "traffic_profile": traffic[1],
                     "kernel_sources": kernel_sources_hash(),
                     "bytes_per_launch": nbytes, "launch_ms": h["dom_ms"],
                     "launches": h["dom_launches"]},
        "cpu_baseline": None if args.no_cpu else cpu_baseline(args.cpu_seconds),
    }
    if not args.no_cpu and not args.headline_only:
        _progress("CPU port leg done")
        ref = reference_algorithm_cpu(progress=_progress)
        device_ms = {
            "c1_bve_gauss12x24_L8": secondary["c1_bve_gauss12x24_L8_fp64_ms_per_step"],
            "c2_bve_geodesic_res4_L21": secondary["geodesic_res4_L21_modes"]["fp64"]["ms_per_step"],
            "c3_swe_gauss128x256_L85_mountain": secondary["swe85_mountain_modes"]["fp64"]["ms_per_step"],
        }
        for name, ms in device_ms.items():
            ref[name]["device_fp64_steps_per_s"] = 1000.0 / ms
            ref[name]["device_over_cpu"] = (1000.0 / ms) / ref[name]["steps_per_s"]
        line["cpu_baseline"]["reference_algorithm"] = ref
    print(json.dumps(line))


if __name__ == "__main__":
    main()
