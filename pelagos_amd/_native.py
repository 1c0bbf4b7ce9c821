"""ctypes binding of the C ABI in include/pelagos_relaybp.h.

The library is built in-tree (``pelagos_amd/libpelagos_relaybp.so``) by
``pelagos_amd/build.py``.  There is no fallback: if the library is missing or
fails to load, :func:`lib` raises, and so does every decoder entry point.
"""

from __future__ import annotations

import ctypes
from pathlib import Path
from typing import Optional

LIB_PATH = Path(__file__).resolve().parent / "libpelagos_relaybp.so"

PEL_ABI_VERSION = 11
PEL_MAX_JOBS = 64
PEL_NOALLOC_MAX_BYTES = 32 << 20
PEL_PLAN_LEARNED = 1
PEL_PLAN_LEGWISE = 2
PEL_MLP_PARAMS = 32 * 8 + 32 + 32 * 32 + 32 + 32 + 1
PEL_OK = 0
PEL_EINVAL = -1
PEL_EDEVICE = -2
PEL_ENOMEM = -3
PEL_EBUSY = -4
PEL_SERVE_WAVE_SHOTS = 64
PEL_DTYPE_F16 = 0
PEL_DTYPE_F32 = 1
PEL_DTYPE_F64 = 2

_vp = ctypes.c_void_p


class PelParams(ctypes.Structure):
    _fields_ = [
        ("pre_iter", ctypes.c_int32),
        ("num_sets", ctypes.c_int32),
        ("set_max_iter", ctypes.c_int32),
        ("stop_nconv", ctypes.c_int32),
        ("plain", ctypes.c_int32),
        ("gamma0", ctypes.c_float),
        ("use_alpha", ctypes.c_int32),
        ("use_beta", ctypes.c_int32),
        ("beta", ctypes.c_float),
        ("alpha_t", _vp),
        ("last_leg", ctypes.c_int32),
        ("skip_relay_legs", ctypes.c_int32),
        ("leg_step", ctypes.c_int32),
    ]


class PelBatch(ctypes.Structure):
    _fields_ = [
        ("B", ctypes.c_int32),
        ("syndromes", _vp),
        ("prior_llr", _vp),
        ("wj", _vp),
        ("prior_batched", ctypes.c_int32),
        ("gamma_legs", _vp),
        ("gamma_per_row", ctypes.c_int32),
        ("gamma_by_position", ctypes.c_int32),
        ("init_gamma", _vp),
        ("init_posterior", _vp),
        ("errors", _vp),
        ("weights", _vp),
        ("iterations", _vp),
        ("found_count", _vp),
        ("last_posterior", _vp),
        ("workspace", _vp),
        ("shot_clock", _vp),
        ("ev_begin", _vp),
        ("ev_end", _vp),
        ("mlp", _vp),
        ("learned_legs", ctypes.c_int32),
        ("feat_out", _vp),
        ("hard_out", _vp),
        ("gamma_out", _vp),
        ("learned_lo", ctypes.c_float),
        ("learned_hi", ctypes.c_float),
        ("learned_last_damp", ctypes.c_float),
    ]


class PelPlan(ctypes.Structure):
    _fields_ = [
        ("block", ctypes.c_int32),
        ("grid", ctypes.c_int32),
        ("placement", ctypes.c_int32),
        ("lds_bytes", ctypes.c_int32),
        ("spill_stride", ctypes.c_int64),
        ("workspace_bytes", ctypes.c_int64),
        ("requeue_below", ctypes.c_int32),
        ("learned", ctypes.c_int32),
        ("legwise", ctypes.c_int32),
    ]


class PelJobsPlan(ctypes.Structure):
    _fields_ = [
        ("n_jobs", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("table_bytes", ctypes.c_int64),
        ("block", ctypes.c_int32),
        ("grid", ctypes.c_int32),
        ("lds_bytes", ctypes.c_int32),
        ("group", ctypes.c_int32 * PEL_MAX_JOBS),
        ("placement", ctypes.c_int32 * PEL_MAX_JOBS),
        ("learned", ctypes.c_int32 * PEL_MAX_JOBS),
        ("job_lds", ctypes.c_int32 * PEL_MAX_JOBS),
        ("spill_stride", ctypes.c_int64 * PEL_MAX_JOBS),
        ("workspace_bytes", ctypes.c_int64 * PEL_MAX_JOBS),
    ]


class PelServeUnit(ctypes.Structure):
    _fields_ = [
        ("job", ctypes.c_int32),
        ("count", ctypes.c_int32),
        ("syn_row_stride", ctypes.c_int64),
        ("syndromes", _vp),
        ("errors", _vp),
        ("weights", _vp),
        ("iterations", _vp),
        ("found_count", _vp),
        ("syn_col_stride", ctypes.c_int64),
    ]


# Every symbol the header declares, with its ctypes signature.
SIGNATURES = {
    "pel_abi_version": (ctypes.c_int, []),
    "pel_last_error": (ctypes.c_char_p, []),
    "pel_graph_create": (ctypes.c_int, [_vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.POINTER(_vp)]),
    "pel_graph_destroy": (ctypes.c_int, [_vp]),
    "pel_graph_info": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64),
                                      ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "pel_graph_permute_vars": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int64, _vp]),
    "pel_plan_decode": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.POINTER(PelPlan)]),
    "pel_decode": (ctypes.c_int, [_vp, ctypes.POINTER(PelParams), ctypes.POINTER(PelBatch),
                                  ctypes.POINTER(PelPlan), _vp]),
    "pel_decode_noalloc": (ctypes.c_int, [_vp, ctypes.POINTER(PelParams), ctypes.POINTER(PelBatch),
                                          ctypes.POINTER(PelPlan), _vp]),
    "pel_plan_jobs": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int32, ctypes.POINTER(PelJobsPlan)]),
    "pel_decode_jobs": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int32, ctypes.POINTER(PelJobsPlan), _vp, _vp]),
    "pel_serve_queue_create": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(_vp)]),
    "pel_serve_queue_destroy": (ctypes.c_int, [_vp]),
    "pel_serve_queue_info": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.POINTER(ctypes.c_int64)),
                                            ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64)]),
    "pel_serve_publish": (ctypes.c_int, [_vp, ctypes.POINTER(PelServeUnit), ctypes.c_int32,
                                         ctypes.POINTER(ctypes.c_int64)]),
    "pel_serve_set_stop": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "pel_serve_done_count": (ctypes.c_int64, [_vp, ctypes.c_int64, ctypes.c_int64]),
    "pel_plan_serve": (ctypes.c_int, [_vp, _vp, ctypes.c_int32, ctypes.POINTER(PelJobsPlan)]),
    "pel_serve_launch": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int32, ctypes.POINTER(PelJobsPlan), _vp, _vp,
                                        ctypes.c_int64, _vp]),
    "pel_plan_typed": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64)]),
    "pel_decode_typed": (ctypes.c_int, [_vp, ctypes.POINTER(PelParams), ctypes.POINTER(PelBatch), ctypes.c_int32,
                                        ctypes.c_int32, _vp]),
    "pel_mlp_param_count": (ctypes.c_int, []),
    "pel_mlp_last_error": (ctypes.c_char_p, []),
    "pel_mlp_grad_workspace_bytes": (ctypes.c_int64, []),
    "pel_mlp_grad": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int64, _vp, _vp, _vp, ctypes.c_int64, _vp]),
    "pel_mlp_adam": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                    ctypes.c_float, ctypes.c_int32, _vp]),
    "pel_mlp_forward": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, _vp, _vp]),
}

_LIB: Optional[ctypes.CDLL] = None


class NativeError(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    """Load libpelagos_relaybp.so (raises if it is not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise NativeError(
            f"HIP decoder library not found at {LIB_PATH}; build it with "
            "`python -m pelagos_amd.build` (the decoder has no CPU fallback)"
        )
    L = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if L.pel_abi_version() != PEL_ABI_VERSION:
        raise NativeError("libpelagos_relaybp.so ABI version mismatch; rebuild it")
    _LIB = L
    return L


def check(rc: int, what: str, mlp: bool = False) -> None:
    """Map a status code to the reference's exception types."""
    if rc == PEL_OK:
        return
    err = lib().pel_mlp_last_error if mlp else lib().pel_last_error
    msg = err().decode(errors="replace")
    if rc == PEL_EINVAL:
        raise ValueError(f"{what}: {msg}")
    if rc == PEL_ENOMEM:
        raise MemoryError(f"{what}: {msg}")
    raise RuntimeError(f"{what}: {msg}")
