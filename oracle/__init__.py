"""CPU oracle for Relay-BP-S decoding (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this package, and only as the checker: the product decoder in
``pelagos_amd`` never calls it and fails loudly when its HIP library is missing.

The C restatement lives in ``relaybp_oracle.c`` (built by ``oracle/Makefile``);
this module binds it with ctypes.  It follows the reference's
``_decode_loop_single_gpu`` (src/pelagos/decoder.py:3394-3873) per shot.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path
from typing import Dict, Optional

import numpy as np

_DIR = Path(__file__).resolve().parent
LIB_PATH = _DIR / "liboracle_relaybp.so"
# the same restatement per element type (relaybp_oracle.c built with ORC_ELEM = 32 / 64 / 16)
LIB_PATHS = {"float32": LIB_PATH, "float64": _DIR / "liboracle_relaybp_f64.so",
             "float16": _DIR / "liboracle_relaybp_f16.so"}
NP_DTYPES = {"float32": np.float32, "float64": np.float64, "float16": np.float16}

_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class _Params(ctypes.Structure):
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
        ("alpha_t", _f32p),
        ("last_leg", ctypes.c_int32),
        ("skip_relay_legs", ctypes.c_int32),
    ]


class _Batch(ctypes.Structure):
    # element-type arrays (priors, strengths, weights, posteriors) as untyped pointers: float,
    # double or half bit patterns, per the library's ORC_ELEM
    _fields_ = [
        ("B", ctypes.c_int32),
        ("syndromes", _u8p),
        ("prior_llr", ctypes.c_void_p),
        ("prior_batched", ctypes.c_int32),
        ("gamma_legs", ctypes.c_void_p),
        ("gamma_per_row", ctypes.c_int32),
        ("init_gamma", ctypes.c_void_p),
        ("init_posterior", ctypes.c_void_p),
        ("errors", _u8p),
        ("weights", ctypes.c_void_p),
        ("iterations", _i32p),
        ("found_count", _i32p),
        ("last_posterior", ctypes.c_void_p),
        ("features0", ctypes.c_void_p),
        ("mu0", ctypes.c_void_p),
        ("nu0", ctypes.c_void_p),
    ]


def build(mutants: bool = False) -> None:
    """Compile the oracle (and optionally its planted-bug mutants) with make."""
    targets = ["all"] + (["mutants"] if mutants else [])
    subprocess.run(["make", "-s", "-C", str(_DIR)] + targets, check=True)


_LIBS: Dict[str, ctypes.CDLL] = {}


def load(path: Optional[os.PathLike] = None) -> ctypes.CDLL:
    """Load (building on first use) the oracle library at ``path``."""
    p = Path(path) if path is not None else LIB_PATH
    key = str(p)
    if key in _LIBS:
        return _LIBS[key]
    if not p.exists():
        build(mutants=p != LIB_PATH)
    lib = ctypes.CDLL(str(p))
    lib.orc_decode.restype = ctypes.c_int
    lib.orc_decode.argtypes = [
        ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _i32p, _i32p,
        ctypes.POINTER(_Params), ctypes.POINTER(_Batch), ctypes.c_int32,
    ]
    if hasattr(lib, "orc_check_update_row"):  # the float32 build
        lib.orc_check_update_row.restype = None
        lib.orc_check_update_row.argtypes = [
            ctypes.c_int32, _f32p, ctypes.c_uint8, ctypes.c_int32, ctypes.c_float,
            ctypes.c_int32, ctypes.c_float, _f32p,
        ]
    if hasattr(lib, "orc_to_half"):  # the float16 build
        for fn in (lib.orc_to_half, lib.orc_from_half):
            fn.restype = None
            fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32]
    lib.orc_planted.restype = ctypes.c_int
    lib.orc_elem_bits.restype = ctypes.c_int
    _LIBS[key] = lib
    return lib


def _ptr(a: Optional[np.ndarray], typ):
    if a is None:
        return ctypes.cast(None, typ)
    return a.ctypes.data_as(typ)


def alpha_schedule(alpha: Optional[float], scaling: float, n: int) -> np.ndarray:
    """Per-iteration alpha (decoder.py:2649-2660), rounded to float32 as a kernel scalar."""
    out = np.ones(max(n, 1), dtype=np.float32)
    if alpha is None:
        return out
    a = float(alpha)
    for t in range(n):
        if a == 0.0:
            s = float(scaling or 1.0)
            out[t] = np.float32(1.0 - (2.0 ** (-((t + 1) / s))))
        elif a < 0.0:
            out[t] = 1.0
        else:
            out[t] = np.float32(a)
    return out


def decode(
    row_ptr: np.ndarray,
    row_var: np.ndarray,
    V: int,
    syndromes: np.ndarray,
    prior_llr: np.ndarray,
    gamma_legs: np.ndarray,
    *,
    pre_iter: int,
    num_sets: int,
    set_max_iter: int,
    stop_nconv: int,
    gamma0: float,
    plain: bool = False,
    alpha: Optional[float] = None,
    alpha_scaling: float = 1.0,
    beta: Optional[float] = None,
    last_leg: bool = False,
    skip_relay_legs: bool = False,
    gamma_per_row: bool = False,
    init_gamma: Optional[np.ndarray] = None,
    init_posterior: Optional[np.ndarray] = None,
    want_posterior: bool = False,
    want_features0: bool = False,
    want_messages0: bool = False,
    nthreads: int = 0,
    lib_path: Optional[os.PathLike] = None,
    dtype: str = "float32",
) -> Dict[str, np.ndarray]:
    """Decode ``syndromes [B,C]`` on the CPU; returns numpy arrays.

    ``gamma_legs`` is the per-leg memory-strength table: ``[num_sets, V]`` or,
    with ``gamma_per_row``, ``[num_sets, B, V]``; row ``l`` is relay leg ``l``.
    ``prior_llr`` is ``[V]`` or per-shot ``[B, V]``.  ``dtype`` is the decoder's
    data_type ("float32", "float64" or "float16"): the element type of the
    priors, strengths, weights and posteriors, which are converted to it first
    (as the reference's decoder casts them, decoder.py:724-734, 3684-3700).
    ``want_features0`` (float32 only) also returns ``features0 [B,V,8]``: the
    learned network's features at the start of relay leg 0 for the shots that
    reach it (zeros for the others), as the decode kernel's feature pass
    computes them.  ``want_messages0`` returns ``mu0`` and ``nu0 [B,E]``: the
    variable-to-check and check-to-variable messages when the pre-iter leg ends
    (edge id = CSR position).
    """
    lib = load(lib_path if lib_path is not None else LIB_PATHS[dtype])
    ft = NP_DTYPES[dtype]
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
    row_var = np.ascontiguousarray(row_var, dtype=np.int32)
    C = row_ptr.shape[0] - 1
    E = int(row_ptr[-1])
    syn = np.ascontiguousarray(syndromes, dtype=np.uint8)
    B = syn.shape[0]
    assert syn.shape == (B, C)
    lam = np.ascontiguousarray(prior_llr, dtype=ft)
    prior_batched = lam.ndim == 2
    assert lam.shape == ((B, V) if prior_batched else (V,))
    g = np.ascontiguousarray(gamma_legs, dtype=ft)
    if num_sets > 0:
        assert g.shape == ((num_sets, B, V) if gamma_per_row else (num_sets, V)), g.shape
    else:
        g = np.zeros((1, V), ft)
    ig = None if init_gamma is None else np.ascontiguousarray(init_gamma, dtype=ft)
    ip = None if init_posterior is None else np.ascontiguousarray(init_posterior, dtype=ft)
    alpha_t = alpha_schedule(alpha, alpha_scaling, max(pre_iter, set_max_iter, 1))

    errors = np.zeros((B, V), np.uint8)
    weights = np.zeros(B, ft)
    iters = np.zeros(B, np.int32)
    found = np.zeros(B, np.int32)
    post = np.zeros((B, V), ft) if want_posterior else None
    if want_features0 and dtype != "float32":
        raise ValueError("want_features0 needs the float32 oracle (the learned form is float32 only)")
    feats = np.zeros((B, V, 8), np.float32) if want_features0 else None
    mu0 = np.zeros((B, max(E, 1)), ft) if want_messages0 else None
    nu0 = np.zeros((B, max(E, 1)), ft) if want_messages0 else None

    prm = _Params(
        pre_iter, num_sets, set_max_iter, stop_nconv, int(bool(plain)), float(gamma0),
        int(alpha is not None), int(beta is not None), float(beta or 0.0),
        _ptr(alpha_t, _f32p), int(bool(last_leg)), int(bool(skip_relay_legs)),
    )
    def vp(a):
        return None if a is None else a.ctypes.data

    bat = _Batch(
        B, _ptr(syn, _u8p), vp(lam), int(prior_batched), vp(g),
        int(bool(gamma_per_row)), vp(ig), vp(ip), _ptr(errors, _u8p),
        vp(weights), _ptr(iters, _i32p), _ptr(found, _i32p), vp(post), vp(feats), vp(mu0), vp(nu0),
    )
    rc = lib.orc_decode(C, V, E, _ptr(row_ptr, _i32p), _ptr(row_var, _i32p),
                        ctypes.byref(prm), ctypes.byref(bat), int(nthreads))
    if rc != 0:
        raise MemoryError("oracle allocation failed")
    out = {
        "errors": errors,
        "weights": weights,
        "iterations": iters,
        "found_count": found,
        "valid_mask": found >= stop_nconv,
    }
    if post is not None:
        out["last_posterior"] = post
    if feats is not None:
        out["features0"] = feats
    if mu0 is not None:
        out["mu0"], out["nu0"] = mu0[:, :E], nu0[:, :E]
    return out


def check_update_row(mu: np.ndarray, syn: int, alpha: Optional[float] = None,
                     beta: Optional[float] = None, lib_path: Optional[os.PathLike] = None) -> np.ndarray:
    """One check's outgoing messages for incoming ``mu`` (kernels.py:1128-1215)."""
    lib = load(lib_path)
    mu = np.ascontiguousarray(mu, dtype=np.float32)
    nu = np.zeros_like(mu)
    lib.orc_check_update_row(mu.shape[0], _ptr(mu, _f32p), int(syn) & 1, int(alpha is not None),
                             float(alpha if alpha is not None else 1.0), int(beta is not None),
                             float(beta or 0.0), _ptr(nu, _f32p))
    return nu


def half_round(x: np.ndarray) -> np.ndarray:
    """The float16 oracle's rounding of doubles to half (bit patterns as np.float16)."""
    lib = load(LIB_PATHS["float16"])
    x = np.ascontiguousarray(x, dtype=np.float64)
    h = np.zeros(x.shape, np.float16)
    lib.orc_to_half(x.ctypes.data, h.ctypes.data, int(x.size))
    return h


def half_value(h: np.ndarray) -> np.ndarray:
    """The float16 oracle's value of half bit patterns, as doubles."""
    lib = load(LIB_PATHS["float16"])
    h = np.ascontiguousarray(h, dtype=np.float16)
    x = np.zeros(h.shape, np.float64)
    lib.orc_from_half(h.ctypes.data, x.ctypes.data, int(h.size))
    return x
