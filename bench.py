"""Benchmark: spectral RK4 steps/s on one MI355X.

Headline workload: barotropic-vorticity RK4 at L = 170 on the Gauss
256 x 512 grid (the L = 170 member class, single member, FP64), fixed
dt = 300 s, seeded l^-1.5 vorticity with rms 5e-5 s^-1.  Also measured in
the same run: the shallow-water step at L = 85 on Gauss 128 x 256 with a
2 km Gaussian mountain (dt = 300 s).  Inputs are resident in HBM before the
timed region; a "step" is one full RK4 step (4 fused stages).

Prints one JSON line.  ``--gpus N`` with N > 1 exits: the measurement
covers one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

R_EARTH = 6.371e6
OMEGA = 2.0 * math.pi / 86400.0
HBM_PEAK_GBS = 8000.0


def _args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=576)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    return p.parse_args()


def _bve_planet(L, nlat, nlon):
    from aeolus_amd.planet import Planet
    return Planet.with_radius(R_EARTH, OMEGA, grid_type="latlon", nlat=nlat, nlon=nlon,
                              l_max=L, product_quadrature="fine")


def _timed_steps(model, y, dt, steps, warmup, prof_cls=None):
    """Run warmup + timed fused steps; return (ms_per_step, y_out, kernel_ms, launches)."""
    import torch
    from aeolus_amd import _native as nat
    out = torch.empty_like(y)
    if warmup:
        model.rk4(y, dt, warmup, out=out)
    torch.cuda.synchronize()
    if prof_cls is not None:
        nat.check(nat.lib().aeo_model_profile(model._model, prof_cls, 4 * steps), "profile")
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record()
    model.rk4(y, dt, steps, out=out)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1)
    kms, nl = None, 0
    if prof_cls is not None:
        import ctypes
        tot = ctypes.c_double()
        cnt = ctypes.c_int()
        nat.check(nat.lib().aeo_model_profile_read(model._model, ctypes.byref(tot),
                                                   ctypes.byref(cnt)), "profile_read")
        nat.check(nat.lib().aeo_model_profile(model._model, -1, 0), "profile off")
        kms, nl = tot.value, cnt.value
    return ms / steps, out, kms, nl


def _class_bytes(cls, L, nlat, nlon, nsyn, nana, nstate):
    """Algorithmic HBM bytes of one launch of a stage kernel class."""
    nh = (nlat + 1) // 2
    B = (L + 1) * (L + 2) // 2
    cut = (2 * L) // 3
    Bc = (cut + 1) * (cut + 2) // 2
    spec = (L + 1) * (L + 1) * 16
    if cls == 0:   # Pbar table + state read + E/O write
        return nh * B * 8 + nstate * spec + 2 * nsyn * nh * (L + 1) * 16
    if cls == 1:   # E/O read + S/D write
        return 2 * nsyn * nh * (L + 1) * 16 + 2 * nana * nh * (cut + 1) * 16
    # analysis: truncated Pbar table + S/D read + stage state r/w (ys, y0, ksum r/w, ynext)
    return nh * Bc * 8 + 2 * nana * nh * (cut + 1) * 16 + 5 * nstate * spec


def run_gpu(args):
    import torch
    from aeolus_amd.physics import BarotropicVorticity, ShallowWaterModel, Topography
    from aeolus_amd.run.initial_conditions import spectrum_l15, swe_spectra

    dev = torch.device("cuda", 0)
    res = {}
    # -- headline: BVE L = 170, Gauss 256 x 512 -------------------------
    L, nlat, nlon = 170, 256, 512
    pl = _bve_planet(L, nlat, nlon)
    model = BarotropicVorticity(pl, viscosity=0.0)
    z0 = torch.as_tensor(spectrum_l15(L, np.random.default_rng(2026), rms=5e-5), device=dev)
    dt = 300.0
    # per-class durations (calibration, untimed)
    cls_ms = {}
    for c in (0, 1, 2):
        _, _, kms, nl = _timed_steps(model, z0, dt, 40, 2, prof_cls=c)
        cls_ms[c] = kms / max(nl, 1)
    dom = max(cls_ms, key=cls_ms.get)
    ms_step, y_out, kms, nl = _timed_steps(model, z0, dt, args.steps, args.warmup, prof_cls=dom)
    ms_plain, _, _, _ = _timed_steps(model, z0, dt, args.steps, 0, prof_cls=None)
    res["bve170"] = dict(ms_per_step=ms_step, ms_per_step_uninstrumented=ms_plain,
                         steps_per_s=1000.0 / ms_step, dom_cls=dom,
                         dom_ms=kms / max(nl, 1), dom_launches=nl,
                         cls_ms={str(k): v for k, v in cls_ms.items()},
                         dom_bytes=_class_bytes(dom, L, nlat, nlon, 4, 1, 1),
                         finite=bool(torch.isfinite(y_out).all()))
    # -- secondary: SWE L = 85, Gauss 128 x 256, mountain ------------------
    L2, nl2, nn2 = 85, 128, 256
    pl2 = _bve_planet(L2, nl2, nn2)
    rng = np.random.default_rng(85)
    lat0 = math.degrees(math.asin(rng.uniform(-1, 1)))
    lon0 = math.degrees(rng.uniform(0, 2 * math.pi))
    topo = Topography.mountain(pl2, height_m=2000.0, lat_deg=lat0, lon_deg=lon0,
                               width_deg=math.degrees(1e6 / R_EARTH))
    swe = ShallowWaterModel(pl2, gravity=9.80616, mean_depth=1e4, topography=topo)
    y0 = torch.as_tensor(swe_spectra(L2, R_EARTH, 9.80616, seed=85), device=dev)
    ms2, y2, _, _ = _timed_steps(swe, y0, 300.0, 288, 10, prof_cls=None)
    res["swe85"] = dict(ms_per_step=ms2, steps_per_s=1000.0 / ms2,
                        finite=bool(torch.isfinite(y2).all()))
    return res, (pl, z0)


def cpu_baseline(seconds: float):
    """Oracle restatement (NumPy FFT + stable Legendre) of the same L = 170 BVE step."""
    from oracle.models import Planet as OPlanet, bve_tendency, rk4_step
    from aeolus_amd.run.initial_conditions import spectrum_l15
    try:
        from threadpoolctl import threadpool_info
        blas = max((i.get("num_threads", 1) for i in threadpool_info()), default=1)
    except Exception:
        blas = 1
    op = OPlanet(l_max=170, radius=R_EARTH, omega=OMEGA, grid_type="latlon", nlat=256,
                 nlon=512)
    y = spectrum_l15(170, np.random.default_rng(2026), rms=5e-5)
    tend = lambda s: bve_tendency(op, s)
    y = rk4_step(tend, y, 300.0)  # warm-up step
    n = 0
    t0 = time.perf_counter()
    while True:
        y = rk4_step(tend, y, 300.0)
        n += 1
        if time.perf_counter() - t0 >= seconds:
            break
    el = time.perf_counter() - t0
    cores = len(os.sched_getaffinity(0))
    return {"value": n / el, "unit": "RK4 steps/s", "cores": int(blas),
            "host_cores_available": cores, "kind": "port",
            "sample": f"{n} RK4 steps of the L=170 Gauss 256x512 BVE step (oracle NumPy "
                      f"FFT + stable-Legendre restatement), after 1 warm-up step"}


def main():
    args = _args()
    if args.gpus != 1:
        print(json.dumps({"error": "measurements cover one GPU; --gpus must be 1"}))
        sys.exit(2)
    res, _ = run_gpu(args)
    b = res["bve170"]
    achieved = b["dom_bytes"] / (b["dom_ms"] * 1e-3) / 1e9
    names = {0: "legendre_syn_kernel<JacLoader> (4 fields)", 1: "grid_jac_kernel (FFT + J)",
             2: "legendre_ana_kernel<EpiBve> (analysis + RK4 epilogue)"}
    line = {
        "metric": "RK4 steps/s at lmax 85/170 vs plain FP64 at rel-err<=1e-11; adjoint:fwd cost",
        "value": b["steps_per_s"],
        "unit": "RK4 steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": b["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded l^-1.5 spectra)",
        "config": {"workload": "BVE RK4, Gauss 256x512, L=170, single member, dt=300 s, "
                               "plain FP64 (split precision not yet built)",
                   "grid": "256x512", "l_max": 170, "dt_s": 300.0, "radius_m": R_EARTH},
        "secondary": {"swe_L85_gauss128x256_mountain_steps_per_s": res["swe85"]["steps_per_s"],
                      "bve170_ms_per_step_uninstrumented": b["ms_per_step_uninstrumented"],
                      "kernel_class_ms": b["cls_ms"]},
        "roofline": {"bound": "hbm", "kernel": names[b["dom_cls"]],
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": b["dom_bytes"], "launch_ms": b["dom_ms"],
                     "launches": b["dom_launches"]},
        "cpu_baseline": None if args.no_cpu else cpu_baseline(args.cpu_seconds),
        "finite": b["finite"] and res["swe85"]["finite"],
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
