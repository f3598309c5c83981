"""Tile-shape autotuner for the Legendre stage kernels.

For a model (a member batch on one Gauss plan) it times a few fused RK4
steps for each candidate number of members per workgroup, keeps the fastest
and records the choice in a JSON cache keyed by device, model kind, L, grid,
member count and precision mode (``$AEOLUS_AMD_CACHE`` or
``~/.cache/aeolus_amd/autotune.json``), so later runs reuse it.
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import torch

__all__ = ["tune_members_per_block", "cache_path"]

CANDIDATES = (1, 2, 4, 8)


def cache_path() -> Path:
    root = os.environ.get("AEOLUS_AMD_CACHE")
    base = Path(root) if root else Path.home() / ".cache" / "aeolus_amd"
    return base / "autotune.json"


def _key(model) -> str:
    dev = torch.cuda.get_device_name(model.device)
    g = model._ps.geometry if model._ps.geometry is not None else model.grid
    kind = type(model).__name__
    nmem = getattr(model, "nmembers", 1)
    return (f"{dev}|{kind}|L{model.sh.l_max}|{g.nlat}x{g.nlon}|n{nmem}|"
            f"p{getattr(model, 'nslices', 0)}")


def _load() -> dict:
    p = cache_path()
    try:
        return json.loads(p.read_text())
    except (OSError, ValueError):
        return {}


def _store(table: dict) -> None:
    p = cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        tmp = p.with_suffix(".tmp")
        tmp.write_text(json.dumps(table, indent=1, sort_keys=True))
        os.replace(tmp, p)
    except OSError:
        pass  # a read-only home only loses the cache, not the tuning


def tune_members_per_block(model, y: torch.Tensor, dt: float, steps: int = 3,
                           candidates=CANDIDATES, use_cache: bool = True) -> dict:
    """Pick the fastest members-per-workgroup tile for ``model``; returns the
    timings (ms per step) and the choice."""
    key = _key(model)
    table = _load() if use_cache else {}
    if use_cache and key in table:
        model.set_tiles(int(table[key]["mb"]))
        return dict(table[key], cached=True)
    nmem = getattr(model, "nmembers", 1)
    out = torch.empty_like(y)
    times = {}
    for mb in candidates:
        if mb > max(1, nmem):
            continue
        try:
            model.set_tiles(mb)
            model.rk4(y, dt, 1, out=out)
            torch.cuda.synchronize()
        except (ValueError, RuntimeError):
            continue  # tile does not fit (LDS) on this size
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record()
        model.rk4(y, dt, steps, out=out)
        t1.record()
        torch.cuda.synchronize()
        times[mb] = t0.elapsed_time(t1) / steps
    if not times:
        raise RuntimeError("autotune: no candidate tile runs on this model")
    best = min(times, key=times.get)
    model.set_tiles(best)
    entry = {"mb": best, "ms_per_step": {str(k): v for k, v in times.items()}}
    if use_cache:
        table[key] = entry
        _store(table)
    return dict(entry, cached=False)
