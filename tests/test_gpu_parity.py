"""HIP decoder vs the CPU oracle, through the C ABI, on seeded inputs.

Bar: bit-exact errors, iterations, found_count and weights, and bit-exact
posteriors (both sides compute the same fp32 operations in the same order).
Cases follow the reference's own tests (tests/test_decoder_fixes.py,
tests/test_unconverged_correction.py, tests/test_triton_against_rust.py)
plus the two decoding workloads: the [[72,12,6]] bivariate bicycle code and
the 936 x 8785 banded circuit-shaped matrix.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
import toy_codes
from pelagos_amd import RelayBPDecoder, codes, utils

pytestmark = pytest.mark.gpu

KEYS = ("errors", "iterations", "found_count", "weights")


def _decoder(H, p, **kw):
    args = dict(pre_iter=80, num_sets=100, set_max_iter=60, gamma0=0.65,
                gamma_dist_interval=(-0.24, 0.66), stop_nconv=5, data_type="float32",
                algo="relay", seed=0, device=[0])
    args.update(kw)
    return RelayBPDecoder(H, p, **args)


def _oracle_for(dec: RelayBPDecoder, syn: np.ndarray, *, prior=None, init_gamma=None,
                init_posterior=None, pre_iter=None, skip_relay_legs=False, want_posterior=True):
    B = syn.shape[0]
    lam = prior if prior is not None else dec.log_prior_ratios_per_gpu[0].reshape(-1).cpu().numpy()
    gtab = dec.gamma_table(B).cpu().numpy()
    return oracle.decode(
        dec.graph.chk_ptr, dec.graph.edge_var, dec.V, syn, lam, gtab,
        pre_iter=dec.pre_iter if pre_iter is None else pre_iter,
        num_sets=dec.num_sets, set_max_iter=dec.set_max_iter, stop_nconv=dec.stop_nconv,
        gamma0=dec.gamma0, plain=dec.algo == "plain", alpha=dec.normalized_min_sum_alpha,
        alpha_scaling=dec.alpha_iteration_scaling_factor, beta=dec.offset_min_sum_beta,
        last_leg=dec.unconverged_correction == "last_leg", skip_relay_legs=skip_relay_legs,
        gamma_per_row=gtab.ndim == 3, init_gamma=init_gamma, init_posterior=init_posterior,
        want_posterior=want_posterior,
    )


def _compare(out, ref, keys=KEYS + ("last_posterior",)):
    for k in keys:
        if k not in ref:
            continue
        got = out[k].cpu().numpy()
        exp = ref[k]
        if not np.array_equal(got, exp):
            bad = np.nonzero((got != exp).reshape(got.shape[0], -1).any(1))[0]
            raise AssertionError(f"{k}: {bad.size} shots differ, first {bad[:8].tolist()}")


def _run(dec, syn, **kw):
    print(f"[gpu] decode B={syn.shape[0]} C={dec.C} V={dec.V} plan={dec.plan(syn.shape[0]).block}",
          flush=True)
    out = dec.decode(torch.from_numpy(syn).to("cuda:0"), return_last_posterior=True, **kw)
    torch.cuda.synchronize()
    return out


def _syndrome_ok(H, errors, syn):
    return (codes.syndromes_of(H, errors) == syn).all(1)


# --- the two decoding workloads ----------------------------------------------------------


def test_bb_code_capacity_matches_oracle():
    """[[72,12,6]] BB code, p = 0.03, 1024 shots, the full Relay-BP-S schedule."""
    H = codes.bb_72_12_6()
    p = np.full(72, 0.03)
    err = codes.sample_errors(p, 1024, seed=0)
    syn = codes.syndromes_of(H, err)
    dec = _decoder(H, p)
    out = _run(dec, syn)
    ref = _oracle_for(dec, syn)
    _compare(out, ref)
    found = out["found_count"].cpu().numpy()
    solved = np.isfinite(out["weights"].cpu().numpy())
    assert _syndrome_ok(H, out["errors"].cpu().numpy()[solved], syn[solved]).all()
    assert np.array_equal(out["valid_mask"].cpu().numpy(), found >= 5)


@pytest.mark.parametrize("placement", [None, 7, 3, 1, 0])
def test_banded_circuit_code_matches_oracle(placement):
    """936 x 8785 banded matrix, log-uniform priors; all-LDS and spilled state."""
    H, p = codes.banded_circuit_code(seed=0)
    err = codes.sample_errors(p, 192, seed=0)
    syn = codes.syndromes_of(H, err)
    dec = _decoder(H, p, state_placement=placement)
    out = _run(dec, syn)
    ref = _oracle_for(dec, syn)
    _compare(out, ref)


@pytest.mark.parametrize("requeue_below", [0, 1000])
def test_requeue_threshold_does_not_change_results(requeue_below):
    """Shots parked between their first leg and their relay legs (every shot whose
    first leg converged, with 1000) decode exactly as shots run in one go (0)."""
    for H, p, B in ((codes.bb_72_12_6(), np.full(72, 0.03), 512),
                    (codes.banded_circuit_code(seed=0)[0], codes.banded_circuit_code(seed=0)[1], 128)):
        err = codes.sample_errors(p, B, seed=5)
        syn = codes.syndromes_of(H, err)
        dec = _decoder(H, p)
        dec.requeue_below = requeue_below
        out = _run(dec, syn)
        assert dec.plan(B).requeue_below == requeue_below
        _compare(out, _oracle_for(dec, syn))


# --- schedule options ------------------------------------------------------------------------


@pytest.fixture(scope="module")
def regular_case():
    """(3,6)-regular code, non-uniform priors, 128 syndromes of increasing difficulty
    (tests/test_decoder_fixes.py:89-97)."""
    H = toy_codes.regular_code(3)
    V = H.shape[1]
    priors = np.random.default_rng(4).uniform(0.01, 0.03, V)
    scale = np.repeat([0.5, 1.0, 1.5, 3.0], 32)[:, None]
    E = (np.random.default_rng(5).random((128, V)) < priors * scale).astype(np.uint8)
    S = codes.syndromes_of(H, E)
    return H, priors, S


OPTION_CASES = {
    "relay_default": dict(pre_iter=10, num_sets=7, set_max_iter=15, stop_nconv=3),
    "plain": dict(algo="plain", pre_iter=30, num_sets=0, set_max_iter=20, stop_nconv=1),
    "alpha_fixed": dict(normalized_min_sum_alpha=0.8, pre_iter=10, num_sets=5, set_max_iter=10),
    "alpha_ramp": dict(normalized_min_sum_alpha=0.0, alpha_iteration_scaling_factor=2.0,
                       pre_iter=10, num_sets=5, set_max_iter=10),
    "beta": dict(offset_min_sum_beta=0.3, pre_iter=10, num_sets=5, set_max_iter=10),
    "last_leg": dict(unconverged_correction="last_leg", pre_iter=5, num_sets=4, set_max_iter=5),
    "no_pre_iter": dict(pre_iter=0, num_sets=6, set_max_iter=12, stop_nconv=2),
    "per_row": dict(gamma_draw="per_row", pre_iter=5, num_sets=8, set_max_iter=10, stop_nconv=2),
    "stop1": dict(pre_iter=30, num_sets=5, set_max_iter=20, stop_nconv=1),
    "zero_legs": dict(pre_iter=30, num_sets=0, set_max_iter=20, stop_nconv=1),
    "set_max_iter_0": dict(pre_iter=200, num_sets=1, set_max_iter=0, stop_nconv=1, gamma0=0.0),
}


@pytest.mark.parametrize("case", sorted(OPTION_CASES))
def test_regular_code_options_match_oracle(regular_case, case):
    H, priors, S = regular_case
    dec = _decoder(H, priors, seed=7, **OPTION_CASES[case])
    out = _run(dec, S)
    ref = _oracle_for(dec, S)
    _compare(out, ref)


@pytest.mark.parametrize("perf", ["default", "rowwise", "throughput", None])
@pytest.mark.parametrize("case", ["relay_default", "last_leg", "per_row"])
def test_both_perf_modes_match_oracle(regular_case, case, perf):
    """The reference runs its parity tests under perf="default" and
    "throughput" (tests/test_decoder_fixes.py:325-341); both modes compute the
    same per-shot results, and so do their aliases here."""
    H, priors, S = regular_case
    dec = _decoder(H, priors, seed=7, perf=perf, **OPTION_CASES[case])
    assert dec.perf == ("throughput" if perf in ("throughput", None) else "default")
    _compare(_run(dec, S), _oracle_for(dec, S))


def test_explicit_gammas_and_by_error_match_oracle(regular_case):
    H, priors, S = regular_case
    V = H.shape[1]
    gammas = np.random.default_rng(6).uniform(-0.24, 0.66, (3, V))
    dec = _decoder(H, priors, pre_iter=10, num_sets=7, set_max_iter=15, stop_nconv=3,
                   explicit_gammas=gammas, seed=7)
    _compare(_run(dec, S), _oracle_for(dec, S))
    intervals = np.stack([np.full(V, -0.3), np.linspace(0.3, 0.8, V)], axis=1)
    dec2 = _decoder(H, priors, pre_iter=5, num_sets=10, set_max_iter=10, stop_nconv=2,
                    gamma_dist_interval_by_error=intervals, seed=7)
    _compare(_run(dec2, S), _oracle_for(dec2, S))


def test_decode_overrides_match_oracle(regular_case):
    """init_gamma, init_posterior, prior_llr, pre_iter_override, skip_relay_legs."""
    H, priors, S = regular_case
    B, V = S.shape[0], H.shape[1]
    rng = np.random.default_rng(11)
    dec = _decoder(H, priors, pre_iter=10, num_sets=6, set_max_iter=10, stop_nconv=2, seed=7)
    ig = rng.uniform(-0.2, 0.7, (B, V)).astype(np.float32)
    ip = (utils.compute_log_prior_ratios(priors)[None, :] * rng.uniform(0.5, 1.5, (B, V))).astype(np.float32)
    pl = utils.compute_log_prior_ratios(priors[None, :] * rng.uniform(0.5, 2.0, (B, V)))
    dev = "cuda:0"
    out = dec.decode(torch.from_numpy(S).to(dev), init_gamma=torch.from_numpy(ig).to(dev),
                     init_posterior=torch.from_numpy(ip).to(dev), prior_llr=torch.from_numpy(pl).to(dev),
                     pre_iter_override=7, return_last_posterior=True)
    ref = _oracle_for(dec, S, prior=pl, init_gamma=ig, init_posterior=ip, pre_iter=7)
    _compare(out, ref)
    out2 = dec.decode(torch.from_numpy(S).to(dev), skip_relay_legs=True, return_last_posterior=True)
    ref2 = _oracle_for(dec, S, skip_relay_legs=True)
    _compare(out2, ref2)


def test_check_magnitudes_clip_at_1e30_like_the_oracle(regular_case):
    """Posteriors beyond 1e30 (an initial posterior of +-1.5e30..5e30 on a third
    of the variables, kept by gamma0 = 1): from the second iteration on the
    incoming |mu| exceed 1e30 and the outgoing magnitudes are clipped there
    (kernels.py:1120-1126), as in the oracle."""
    H, priors, S = regular_case
    B, V = S.shape[0], H.shape[1]
    rng = np.random.default_rng(13)
    lam = utils.compute_log_prior_ratios(priors)
    ip = np.broadcast_to(lam, (B, V)).astype(np.float32).copy()
    big = rng.random((B, V)) < 0.33
    ip[big] = (np.where(rng.random(big.sum()) < 0.5, -1.0, 1.0) * rng.uniform(1.5e30, 5e30, big.sum())).astype(np.float32)
    dec = _decoder(H, priors, pre_iter=12, num_sets=4, set_max_iter=10, stop_nconv=2, gamma0=1.0, seed=7)
    out = dec.decode(torch.from_numpy(S).to("cuda:0"), init_posterior=torch.from_numpy(ip).to("cuda:0"),
                     return_last_posterior=True)
    torch.cuda.synchronize()
    _compare(out, _oracle_for(dec, S, init_posterior=ip))


def test_toy_codes_match_oracle():
    """Triangle, high-degree, a 100-wide check, degree-1 and empty checks."""
    dense = np.zeros((5, 120), np.uint8)
    dense[0, :100] = 1          # 100-wide row: four sign words
    dense[1, 100] = 1           # degree-1 check
    dense[3, [5, 50, 110]] = 1  # row 2 stays empty
    dense[4, 100:120] = 1
    from scipy.sparse import csr_matrix

    cases = [
        (toy_codes.triangle_code(), 0.01),
        (toy_codes.high_degree_code(), 0.01),
        (csr_matrix(dense), 0.02),
        (toy_codes.repetition_code(9), 0.1),
    ]
    for H, pval in cases:
        V = H.shape[1]
        p = np.full(V, pval)
        err = codes.sample_errors(p, 64, seed=3, scale=3.0)
        syn = codes.syndromes_of(H, err)
        dec = _decoder(H, p, pre_iter=20, num_sets=5, set_max_iter=10, stop_nconv=2)
        _compare(_run(dec, syn), _oracle_for(dec, syn))


def test_block_sizes_match_oracle():
    """The same decode with every workgroup size."""
    H, p = codes.banded_circuit_code(n_blocks=4, V=2700, seed=2)
    err = codes.sample_errors(p, 64, seed=1, scale=2.0)
    syn = codes.syndromes_of(H, err)
    ref = None
    for block in (64, 128, 256, 512, 1024):
        dec = _decoder(H, p, block_size=block)
        out = _run(dec, syn)
        if ref is None:
            ref = _oracle_for(dec, syn)
        _compare(out, ref)


# --- invariants of the reference tests -------------------------------------------------------


def test_batch_subset_and_single_shot_invariance(regular_case):
    """A shot decodes the same alone, in a subset, and in the full batch
    (tests/test_decoder_fixes.py:184-268)."""
    H, priors, S = regular_case
    dec = _decoder(H, priors, pre_iter=10, num_sets=20, set_max_iter=20, stop_nconv=2, seed=7)
    full = _run(dec, S)
    rows = np.random.default_rng(4).permutation(S.shape[0])[:10]
    part = _run(dec, S[rows])
    for k in ("errors", "found_count", "iterations", "weights"):
        assert torch.equal(full[k][torch.from_numpy(rows).to(full[k].device)], part[k]), k
    one = _run(dec, S[rows[:1]])
    for k in ("errors", "found_count", "iterations"):
        assert torch.equal(full[k][int(rows[0])], one[k][0]), k


def test_empty_and_all_zero_syndrome_batches(regular_case):
    """An empty batch returns empty results without a launch; an all-zero
    syndrome batch decodes to the zero error at weight 0, as the oracle does."""
    H, priors, S = regular_case
    dec = _decoder(H, priors, pre_iter=10, num_sets=20, set_max_iter=20, stop_nconv=2, seed=7)
    out = _run(dec, S[:0])
    assert tuple(out["errors"].shape) == (0, dec.V) and out["errors"].dtype == torch.uint8
    for k in ("weights", "iterations", "found_count", "valid_mask", "unconverged_indices"):
        assert out[k].shape[0] == 0, k
    zero = np.zeros((5, dec.C), np.uint8)
    out = _run(dec, zero)
    assert not out["errors"].any() and (out["weights"] == 0).all()
    assert out["valid_mask"].all()
    _compare(out, _oracle_for(dec, zero))


def test_never_converged_counts_every_iteration():
    """A syndrome no error pattern explains never converges and reports
    pre_iter + num_sets * set_max_iter (tests/test_decoder_fixes.py:367-374)."""
    dense = np.zeros((3, 6), np.uint8)
    dense[0, :3] = 1
    dense[1, 3:] = 1      # row 2 is empty: syndrome bit 1 there is unsatisfiable
    from scipy.sparse import csr_matrix

    H = csr_matrix(dense)
    syn = np.array([[1, 0, 1], [0, 1, 1]], np.uint8)
    dec = _decoder(H, np.full(6, 0.05), pre_iter=30, num_sets=5, set_max_iter=20, stop_nconv=1)
    out = _run(dec, syn)
    assert out["iterations"].cpu().tolist() == [130, 130]
    assert out["found_count"].cpu().tolist() == [0, 0]
    assert torch.isinf(out["weights"]).all()
    _compare(out, _oracle_for(dec, syn))


# --- fast_max_unconverged_pct: the leg-synchronous early exit ---------------------------------


def _unconverged(out, S):
    return float((out["found_count"] < S).float().mean())


@pytest.mark.parametrize("correction", ["first_leg", "last_leg"])
def test_early_exit_after_relay_leg_one_equals_a_one_leg_decode(regular_case, correction):
    """With the threshold between the unconverged fractions after the pre-iter
    leg and after relay leg 1, the decode exits after relay leg 1: its results
    equal a decode with num_sets = 1 (and the oracle's), bit for bit
    (tests/test_unconverged_correction.py:105-121, 177-194)."""
    H, priors, S = regular_case
    kw = dict(pre_iter=2, set_max_iter=15, stop_nconv=1, seed=7, unconverged_correction=correction)
    dev = "cuda:0"
    syn = torch.from_numpy(S).to(dev)
    leg0 = _decoder(H, priors, num_sets=0, **kw).decode(syn)
    one = _decoder(H, priors, num_sets=1, **kw)
    leg1 = one.decode(syn, return_last_posterior=True)
    u0, u1 = _unconverged(leg0, 1), _unconverged(leg1, 1)
    assert u1 < u0 < 1.0, (u0, u1)
    dec = _decoder(H, priors, num_sets=6, fast_max_unconverged_pct=(u0 + u1) / 2, **kw)
    out = dec.decode(syn, return_last_posterior=True)
    torch.cuda.synchronize()
    assert len(out["unconverged_indices"]) > 0
    for k in ("errors", "weights", "iterations", "found_count", "last_posterior", "valid_mask"):
        assert torch.equal(out[k], leg1[k]), k
    _compare(out, _oracle_for(one, S))


def test_early_exit_threshold_zero_runs_the_whole_schedule(regular_case):
    """A threshold of 0 never exits early: the leg-synchronous decode equals
    the whole-schedule decode, on the plain and the learned form."""
    from pelagos_amd import learned

    H, priors, S = regular_case
    syn = torch.from_numpy(S).to("cuda:0")
    for extra in ({}, dict(learned_gamma=learned.GammaMLP(seed=3), learned_legs=2)):
        kw = dict(pre_iter=8, num_sets=6, set_max_iter=8, stop_nconv=3, seed=7, **extra)
        whole = _decoder(H, priors, **kw).decode(syn, return_last_posterior=True)
        stepped = _decoder(H, priors, fast_max_unconverged_pct=0.0, **kw).decode(syn, return_last_posterior=True)
        torch.cuda.synchronize()
        for k in ("errors", "weights", "iterations", "found_count", "last_posterior"):
            assert torch.equal(stepped[k], whole[k]), (bool(extra), k)


def test_bitpack_output_and_llr_column_overrides(regular_case):
    """bitpack_output=True returns the same errors packed 32 to an int32 word
    (utils.py:32-60); override_llr_columns writes priors and selection weights
    of some columns past the clip, the decode follows them as the oracle does
    with those priors, and the returned callable restores the decoder
    (decoder.py:786-835)."""
    from pelagos_amd import utils

    H, priors, S = regular_case
    kw = dict(pre_iter=8, num_sets=6, set_max_iter=8, stop_nconv=3, seed=7)
    dec = _decoder(H, priors, **kw)
    packed = _decoder(H, priors, bitpack_output=True, **kw)
    syn = torch.from_numpy(S).to("cuda:0")
    print("[gpu] bitpack and overrides", flush=True)
    base = dec.decode(syn)
    outp = packed.decode(syn)
    torch.cuda.synchronize()
    V = H.shape[1]
    assert outp["errors"].dtype == torch.int32 and tuple(outp["errors"].shape) == (S.shape[0], (V + 31) // 32)
    assert torch.equal(outp["errors"], utils.bitpack_errors(base["errors"]))
    assert torch.equal(utils.bitunpack_errors(outp["errors"], V), base["errors"])
    for k in ("weights", "iterations", "found_count"):
        assert torch.equal(outp[k], base[k]), k

    undo_a = dec.override_llr_columns([0, 5, 17], 25.0)
    undo_b = dec.override_llr_columns([40, 41], -2.0)
    out = _run(dec, S)
    _compare(out, _oracle_for(dec, S))    # the decoder's priors now hold the overrides
    assert not torch.equal(out["errors"], base["errors"]) or not torch.equal(out["weights"], base["weights"])
    undo_b()
    undo_a()
    again = dec.decode(syn)
    torch.cuda.synchronize()
    for k in ("errors", "weights", "iterations", "found_count"):
        assert torch.equal(again[k], base[k]), k
