"""Tile-shape autotuner for the Legendre stage kernels.

For a model (a member batch on one Gauss plan) it times a few fused RK4
steps for each candidate tile shape, keeps the fastest and records the
choice in a JSON cache keyed by device, model kind, L, grid, member count
and precision mode (``$AEOLUS_AMD_CACHE`` or
``~/.cache/aeolus_amd/autotune.json``), so later runs reuse it.

Candidates: the fused stage kernel with 1, 2, 4 or 8 members per workgroup,
and, for split-precision BVE batches, the GEMM form (``members_per_block``
0) with 64 or 128 synthesis columns per workgroup, 4 or 8 members per
epilogue workgroup and 2 or 4 MFMA column tiles per synthesis wave.  Every
candidate gives bitwise the same result.
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import torch

__all__ = ["tune_members_per_block", "cache_path"]

CANDIDATES = (1, 2, 4, 8)
# (syn_cols, epi_members, syn_ntiles)
GEMM_CANDIDATES = tuple((c, e, t) for c in (64, 128) for e in (4, 8) for t in (2, 4))


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


def _label(mb, gemm) -> str:
    return f"mb{mb}" if gemm is None else "gemm" + "x".join(str(v) for v in gemm)


def _apply(model, mb, gemm) -> None:
    model.set_tiles(mb)
    if gemm is not None:
        model.set_gemm_tiles(*gemm)


def tune_members_per_block(model, y: torch.Tensor, dt: float, steps: int = 3,
                           candidates=CANDIDATES, use_cache: bool = True,
                           gemm_candidates=GEMM_CANDIDATES) -> dict:
    """Pick the fastest tile shape for ``model``; returns the timings (ms per
    step, by label ``mb<k>`` or ``gemm<cols>x<members>x<tiles>``) and the choice
    (``mb``, and ``gemm`` = [syn_cols, epi_members, syn_ntiles] when ``mb``
    is 0)."""
    key = _key(model)
    table = _load() if use_cache else {}
    if use_cache and key in table:
        e = table[key]
        _apply(model, int(e["mb"]), tuple(e["gemm"]) if e.get("gemm") else None)
        return dict(e, cached=True)
    nmem = getattr(model, "nmembers", 1)
    cands = [(mb, None) for mb in candidates if mb <= max(1, nmem)]
    if getattr(model, "nslices", 0) and hasattr(model, "set_gemm_tiles"):
        cands += [(0, g) for g in gemm_candidates]
    out = torch.empty_like(y)
    times = {}
    for mb, gemm in cands:
        try:
            _apply(model, mb, gemm)
            model.rk4(y, dt, 1, out=out)
            torch.cuda.synchronize()
        except (ValueError, RuntimeError):
            continue  # tile does not fit (LDS) or does not apply to this model
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record()
        model.rk4(y, dt, steps, out=out)
        t1.record()
        torch.cuda.synchronize()
        times[(mb, gemm)] = t0.elapsed_time(t1) / steps
    if not times:
        raise RuntimeError("autotune: no candidate tile runs on this model")
    best_mb, best_gemm = min(times, key=times.get)
    _apply(model, best_mb, best_gemm)
    entry = {"mb": best_mb, "gemm": list(best_gemm) if best_gemm else None,
             "ms_per_step": {_label(*k): v for k, v in times.items()}}
    if use_cache:
        table[key] = entry
        _store(table)
    return dict(entry, cached=False)
