"""data_type "float64" and "float16" on the GPU (pel_decode_typed) against the
oracle's restatement in the same element type, bit for bit: errors,
iterations, found counts, weights and last posteriors.  Also: float32 through
the same per-edge kernel equals the fast float32 kernel bit for bit.

The reference accepts these data types (decoder.py:402-423) and runs its
kernels' dtype_enum 2 / 0 branches; its own tests run the dtype comparison
against relay-bp (tests/test_triton_against_rust.py:708-739), which is not
available here.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import oracle
import toy_codes
from pelagos_amd import RelayBPDecoder, _native, codes, utils

pytestmark = pytest.mark.gpu

KEYS = ("errors", "iterations", "found_count", "weights", "last_posterior")
# "float32-typed": a float32 decoder routed through the per-edge kernel (pel_decode_typed)
DTYPES = ("float64", "float16", "float32-typed")


def _decoder(H, p, dtype, **kw):
    args = dict(pre_iter=80, num_sets=100, set_max_iter=60, gamma0=0.65,
                gamma_dist_interval=(-0.24, 0.66), stop_nconv=5, data_type=dtype.split("-")[0],
                algo="relay", seed=0, device=[0])
    args.update(kw)
    dec = RelayBPDecoder(H, p, **args)
    if dtype == "float32-typed":
        dec._typed = _native.PEL_DTYPE_F32
    return dec


def _oracle_for(dec, syn, *, prior=None, init_posterior=None, init_gamma=None, skip_relay_legs=False):
    B = syn.shape[0]
    lam = prior if prior is not None else dec.log_prior_ratios_per_gpu[0].reshape(-1).cpu().numpy()
    gtab = dec.gamma_table(B).cpu().numpy()
    return oracle.decode(
        dec.graph.chk_ptr, dec.graph.edge_var, dec.V, syn, lam, gtab,
        pre_iter=dec.pre_iter, num_sets=dec.num_sets, set_max_iter=dec.set_max_iter,
        stop_nconv=dec.stop_nconv, gamma0=dec.gamma0, plain=dec.algo == "plain",
        alpha=dec.normalized_min_sum_alpha, alpha_scaling=dec.alpha_iteration_scaling_factor,
        beta=dec.offset_min_sum_beta, last_leg=dec.unconverged_correction == "last_leg",
        skip_relay_legs=skip_relay_legs, gamma_per_row=gtab.ndim == 3, init_posterior=init_posterior,
        init_gamma=init_gamma,
        want_posterior=True, dtype=dec.dtype_str,
    )


def _run(dec, syn, **kw):
    print(f"[gpu] {dec.dtype_str} decode B={syn.shape[0]} C={dec.C} V={dec.V}", flush=True)
    out = dec.decode(torch.from_numpy(syn).to("cuda:0"), return_last_posterior=True, **kw)
    torch.cuda.synchronize()
    return out


def _compare(out, ref):
    for k in KEYS:
        got = out[k].cpu().numpy()
        exp = ref[k]
        assert got.dtype == exp.dtype, (k, got.dtype, exp.dtype)
        if not np.array_equal(got, exp, equal_nan=True):
            bad = np.nonzero((got != exp).reshape(got.shape[0], -1).any(1))[0]
            raise AssertionError(f"{k}: {bad.size} shots differ, first {bad[:8].tolist()}")


@pytest.mark.parametrize("dtype", DTYPES)
def test_bb_code_capacity_matches_oracle(dtype):
    H = codes.bb_72_12_6()
    p = np.full(72, 0.03)
    syn = codes.syndromes_of(H, codes.sample_errors(p, 512, seed=0))
    dec = _decoder(H, p, dtype)
    out = _run(dec, syn)
    _compare(out, _oracle_for(dec, syn))
    solved = np.isfinite(out["weights"].cpu().numpy())
    assert (codes.syndromes_of(H, out["errors"].cpu().numpy()[solved]) == syn[solved]).all()


@pytest.mark.parametrize("dtype", DTYPES)
def test_banded_circuit_code_matches_oracle(dtype):
    H, p = codes.banded_circuit_code(seed=0)
    syn = codes.syndromes_of(H, codes.sample_errors(p, 48, seed=0))
    dec = _decoder(H, p, dtype)
    _compare(_run(dec, syn), _oracle_for(dec, syn))


CASES = {
    "plain": dict(algo="plain", pre_iter=30, num_sets=0, set_max_iter=20, stop_nconv=1),
    "alpha": dict(normalized_min_sum_alpha=0.8, pre_iter=10, num_sets=5, set_max_iter=10),
    "alpha_ramp": dict(normalized_min_sum_alpha=0.0, alpha_iteration_scaling_factor=2.0, pre_iter=10,
                       num_sets=5, set_max_iter=10),
    "beta": dict(offset_min_sum_beta=0.3, pre_iter=10, num_sets=5, set_max_iter=10),
    "last_leg": dict(unconverged_correction="last_leg", pre_iter=5, num_sets=4, set_max_iter=5),
    "no_pre_iter": dict(pre_iter=0, num_sets=6, set_max_iter=12, stop_nconv=2),
    "per_row": dict(gamma_draw="per_row", pre_iter=5, num_sets=8, set_max_iter=10, stop_nconv=2),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("dtype", DTYPES)
def test_regular_code_options_match_oracle(dtype, case):
    H = toy_codes.regular_code(3)
    V = H.shape[1]
    priors = np.random.default_rng(4).uniform(0.01, 0.03, V)
    scale = np.repeat([0.5, 1.0, 2.0, 3.0], 16)[:, None]
    E = (np.random.default_rng(5).random((64, V)) < priors * scale).astype(np.uint8)
    S = codes.syndromes_of(H, E)
    dec = _decoder(H, priors, dtype, seed=7, **CASES[case])
    _compare(_run(dec, S), _oracle_for(dec, S))


@pytest.mark.parametrize("dtype", DTYPES)
def test_overrides_match_oracle(dtype):
    """prior_llr, init_posterior, init_gamma and skip_relay_legs in the element type."""
    H = toy_codes.regular_code(3)
    V = H.shape[1]
    priors = np.random.default_rng(4).uniform(0.01, 0.03, V)
    E = (np.random.default_rng(6).random((32, V)) < priors * 2).astype(np.uint8)
    S = codes.syndromes_of(H, E)
    rng = np.random.default_rng(11)
    T = {"float64": np.float64, "float16": np.float16, "float32-typed": np.float32}[dtype]
    ip = (utils.compute_log_prior_ratios(priors)[None, :] * rng.uniform(0.5, 1.5, (32, V))).astype(T)
    pl = utils.compute_log_prior_ratios(priors[None, :] * rng.uniform(0.5, 2.0, (32, V))).astype(T)
    dec = _decoder(H, priors, dtype, pre_iter=10, num_sets=6, set_max_iter=10, stop_nconv=2, seed=7)
    out = _run(dec, S, init_posterior=torch.from_numpy(ip).to("cuda:0"), prior_llr=torch.from_numpy(pl).to("cuda:0"))
    _compare(out, _oracle_for(dec, S, prior=pl, init_posterior=ip))
    out2 = _run(dec, S, skip_relay_legs=True)
    _compare(out2, _oracle_for(dec, S, skip_relay_legs=True))
    ig = rng.uniform(-0.24, 0.66, (32, V)).astype(T)   # the pre-iter leg's per-variable strengths
    out3 = _run(dec, S, init_gamma=torch.from_numpy(ig).to("cuda:0"))
    _compare(out3, _oracle_for(dec, S, init_gamma=ig))


def test_float32_per_edge_kernel_equals_the_fast_kernel():
    """pel_decode_typed with PEL_DTYPE_F32 (the per-edge formulation) gives the
    fast kernels' results bit for bit: the same float32 semantics."""
    H, p = codes.banded_circuit_code(seed=0)
    syn_np = codes.syndromes_of(H, codes.sample_errors(p, 64, seed=2, scale=5.0))
    dec = _decoder(H, p, "float32")
    syn = torch.from_numpy(syn_np).to("cuda:0").contiguous()  # its buffer goes to the C ABI as is
    fast = dec.decode(syn, return_last_posterior=True)
    B, V = syn_np.shape[0], dec.V
    lib = _native.lib()
    grid = ctypes.c_int32()
    nbytes = ctypes.c_int64()
    _native.check(lib.pel_plan_typed(dec._handle.h, B, _native.PEL_DTYPE_F32, 0, ctypes.byref(grid),
                                     ctypes.byref(nbytes)), "pel_plan_typed")
    ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device="cuda:0")
    out = {k: torch.empty_like(fast[k]) for k in ("errors", "weights", "iterations", "found_count", "last_posterior")}
    params, batch = dec._args(syn, dec.log_prior_ratios_per_gpu[0].reshape(V), dec.wj, False, dec.gamma_table(B),
                              None, None, out["errors"], out["weights"], out["iterations"], out["found_count"],
                              out["last_posterior"], dec.pre_iter, False, B, None, ws, by_position=False)
    stream = torch.cuda.current_stream().cuda_stream
    print(f"[gpu] pel_decode_typed float32 B={B} grid={grid.value}", flush=True)
    _native.check(lib.pel_decode_typed(dec._handle.h, ctypes.byref(params), ctypes.byref(batch),
                                       _native.PEL_DTYPE_F32, grid.value, ctypes.c_void_p(stream)), "pel_decode_typed")
    torch.cuda.synchronize()
    for k in out:
        assert torch.equal(out[k], fast[k]), k


def test_typed_decoders_refuse_float32_only_features():
    H = codes.bb_72_12_6()
    p = np.full(72, 0.03)
    dec = _decoder(H, p, "float64")
    with pytest.raises(NotImplementedError):
        dec.prepare(16)


# --- fast_max_unconverged_pct: leg-stepped launches of the per-edge kernel ------------------


def _regular_case():
    H = toy_codes.regular_code(3)
    V = H.shape[1]
    priors = np.random.default_rng(4).uniform(0.01, 0.03, V)
    E = (np.random.default_rng(8).random((96, V)) < priors * np.repeat([1.0, 2.0, 3.0], 32)[:, None])
    return H, priors, codes.syndromes_of(H, E.astype(np.uint8))


def _unconverged(out, S):
    return float((out["found_count"] < S).float().mean())


@pytest.mark.parametrize("correction", ["first_leg", "last_leg"])
@pytest.mark.parametrize("dtype", DTYPES)
def test_early_exit_after_relay_leg_one_equals_a_one_leg_decode(dtype, correction):
    """With the threshold between the unconverged fractions after the pre-iter
    leg and after relay leg 1, the leg-stepped decode exits after relay leg 1:
    its results equal a decode with num_sets = 1 and the oracle's, bit for bit
    (tests/test_unconverged_correction.py:105-121, 177-194)."""
    H, priors, S = _regular_case()
    kw = dict(pre_iter=2, set_max_iter=15, stop_nconv=1, seed=7, unconverged_correction=correction)
    leg0 = _run(_decoder(H, priors, dtype, num_sets=0, **kw), S)
    one = _decoder(H, priors, dtype, num_sets=1, **kw)
    leg1 = _run(one, S)
    u0, u1 = _unconverged(leg0, 1), _unconverged(leg1, 1)
    assert u1 < u0 < 1.0, (u0, u1)
    out = _run(_decoder(H, priors, dtype, num_sets=6, fast_max_unconverged_pct=(u0 + u1) / 2, **kw), S)
    assert len(out["unconverged_indices"]) > 0
    for k in KEYS + ("valid_mask",):
        assert torch.equal(out[k], leg1[k]), k
    _compare(out, _oracle_for(one, S))


@pytest.mark.parametrize("dtype", DTYPES)
def test_early_exit_threshold_zero_runs_the_whole_schedule(dtype):
    """A threshold of 0 never exits early: the leg-stepped decode equals the
    whole-schedule decode and the oracle, bit for bit."""
    H, priors, S = _regular_case()
    kw = dict(pre_iter=8, num_sets=6, set_max_iter=8, stop_nconv=3, seed=7)
    whole_dec = _decoder(H, priors, dtype, **kw)
    whole = _run(whole_dec, S)
    stepped = _run(_decoder(H, priors, dtype, fast_max_unconverged_pct=0.0, **kw), S)
    for k in KEYS:
        assert torch.equal(stepped[k], whole[k]), k
    _compare(stepped, _oracle_for(whole_dec, S))
