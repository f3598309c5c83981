"""ctypes binding of ``libaeolus_hip.so`` (the C ABI declared in include/aeolus_hip.h).

The library is the only compute path: importing a numerical class without
the built library raises immediately; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import torch

_LIB_PATH = Path(__file__).resolve().parent / "lib" / "libaeolus_hip.so"

AEO_OK = 0
AEO_ERR_INVALID = -1
AEO_ERR_UNSUPPORTED = -2
AEO_ERR_HIP = -3
AEO_ERR_STATE = -4

OP_LAP, OP_INVLAP, OP_DLAMBDA, OP_SINTH_DTH, OP_SINTH_DTH_T, OP_TRUNC23, OP_INVLAP_ZERO = range(7)
GRID_STATS_BVE, GRID_STATS_SWE, GRID_STATS_MINMAX = range(3)

_c_int = ctypes.c_int
_c_double = ctypes.c_double
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)

# name -> (argtypes)
_SIGNATURES = {
    "aeo_version": [],
    "aeo_last_error": [],
    "aeo_plan_create_gauss": [_c_int, _c_int, _c_int, _dp, _dp, _dp, _c_int, ctypes.POINTER(_vp)],
    "aeo_plan_create_points": [_c_int, _c_int, _dp, _dp, _dp, _dp, _c_int, ctypes.POINTER(_vp)],
    "aeo_plan_destroy": [_vp],
    "aeo_plan_enable_split": [_vp, _c_int],
    "aeo_model_set_precision": [_vp, _c_int],
    "aeo_model_set_tiles": [_vp, _c_int],
    "aeo_model_set_gemm_tiles": [_vp, _c_int, _c_int, _c_int],
    "aeo_plan_info": [_vp, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int), ctypes.POINTER(_c_int)],
    "aeo_plan_basis": [_vp, _vp, _vp],
    "aeo_analysis": [_vp, _vp, _c_int, _vp, _vp],
    "aeo_synthesis": [_vp, _vp, _c_int, _vp, _vp],
    "aeo_spec_op": [_c_int, _c_int, _c_double, _vp, _vp, _c_int, _vp],
    "aeo_jacobian": [_vp, _c_double, _vp, _vp, _vp, _c_int, _vp],
    "aeo_velocity": [_vp, _c_double, _vp, _vp, _vp, _vp],
    "aeo_gradient": [_vp, _c_double, _vp, _vp, _vp, _vp],
    "aeo_bve_create": [_vp, _c_double, _c_double, _c_double, ctypes.POINTER(_vp)],
    "aeo_bve_create_batch": [_vp, _c_double, _c_double, _c_double, _c_int, ctypes.POINTER(_vp)],
    "aeo_swe_create": [_vp, _c_double, _c_double, _c_double, _c_double, _c_double, _vp,
                       ctypes.POINTER(_vp)],
    "aeo_model_destroy": [_vp],
    "aeo_model_tendency": [_vp, _vp, _vp, _vp, _vp],
    "aeo_model_rk4": [_vp, _vp, _c_double, _c_int, _vp, _vp, _vp],
    "aeo_model_enable_adjoint": [_vp],
    "aeo_model_checkpoint_elems": [_vp, _c_int, ctypes.POINTER(ctypes.c_longlong)],
    "aeo_model_rk4_record": [_vp, _vp, _c_double, _c_int, _vp, _vp, _vp, _vp],
    "aeo_model_rk4_adjoint": [_vp, _vp, _c_double, _c_int, _vp, _vp, _vp],
    "aeo_model_tendency_adjoint": [_vp, _vp, _vp, _vp, _vp],
    "aeo_spectral_diag": [_c_int, _c_double, _c_double, _vp, _c_int, _vp, _vp],
    "aeo_grid_stats": [_c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "aeo_model_profile": [_vp, _c_int, _c_int],
    "aeo_ensemble_create": [_c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp)],
    "aeo_ensemble_destroy": [_vp],
    "aeo_ensemble_rk4": [_vp, ctypes.POINTER(_vp), _c_double, _c_int, ctypes.POINTER(_vp), _c_int,
                         _vp],
    "aeo_model_profile_read": [_vp, _dp, ctypes.POINTER(_c_int)],
    "aeo_vector_curl_div_work": [_vp, _c_int, ctypes.POINTER(ctypes.c_longlong)],
    "aeo_vector_curl_div": [_vp, _c_double, _vp, _vp, _c_int, _c_int, _vp, _vp, _vp, _vp],
    "aeo_vector_curl_div_noalloc": [_vp, _c_double, _vp, _vp, _c_int, _c_int, _vp, _vp, _vp],
    "aeo_pe_create": [_vp, _vp, _c_int, _dp, _c_double, _c_double, _c_double, _c_double, _vp,
                      ctypes.POINTER(_vp)],
    "aeo_pe_destroy": [_vp],
    "aeo_pe_tendency": [_vp, _vp, _vp, _vp],
}

_lib = None


class NativeError(RuntimeError):
    """A HIP runtime failure reported by the library."""


def library_path() -> Path:
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    """Load (once) and return the HIP library; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise ImportError(
            f"aeolus_amd: HIP library not found at {_LIB_PATH}; build it with "
            f"`python -m aeolus_amd.build` (there is no CPU fallback)")
    L = ctypes.CDLL(str(_LIB_PATH))
    for name, args in _SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_char_p if name == "aeo_last_error" else _c_int
    _lib = L
    return L


def check(rc: int, what: str) -> None:
    """Map a status code to the reference's exception conventions."""
    if rc == AEO_OK:
        return
    msg = lib().aeo_last_error().decode(errors="replace")
    if rc == AEO_ERR_UNSUPPORTED:
        raise ValueError(f"{what}: {msg} (no silent fallback)")
    if rc == AEO_ERR_INVALID:
        raise ValueError(f"{what}: {msg}")
    raise NativeError(f"{what}: {msg} (status {rc})")


def host_ptr(arr) -> ctypes.POINTER(ctypes.c_double):
    """Pointer to a C-contiguous float64 NumPy array (kept alive by the caller)."""
    return arr.ctypes.data_as(_dp)


def dptr(t: torch.Tensor | None) -> int | None:
    """Device pointer of a CUDA tensor (None passes NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("aeolus_amd: expected a device (CUDA/HIP) tensor")
    if not t.is_contiguous():
        raise ValueError("aeolus_amd: expected a contiguous tensor")
    return t.data_ptr()


def stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def default_device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("aeolus_amd: no HIP device is visible (the library has no CPU path)")
    return torch.device("cuda", torch.cuda.current_device())

