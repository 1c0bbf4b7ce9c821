"""Pins for the oracle's float64 and float16 restatements (relaybp_oracle.c built
with ORC_ELEM = 64 / 16), independent of the oracle's code.

The reference runs data_type "float64" / "float16" through its kernels'
dtype_enum 2 / 0 branches (kernels.py:1115-1126, 1299-1350, 676-715): float64
stores and computes every value in double; float16 stores messages, posteriors,
priors and strengths as half, takes check minima and variable sums in fp32 and
the memory mix in half arithmetic.  These checks fix:

* the half conversions: every one of the 65536 bit patterns converts to its
  value, and doubles round to the nearest half with ties to even, subnormals,
  overflow and signed zeros as numpy's float16 does;
* the first iteration and the memory mix against closed forms computed here
  with numpy in the element type (double, or half operations);
* tree codes: min-sum in either type still returns the brute-force
  maximum-likelihood pattern;
* every claimed solution satisfies H e = s, with the weight its bits sum to.

Bit-level agreement with the reference's Triton fp16 kernels (their exact
rounding points and any contraction) needs Triton on a GPU and is unpinned.
"""

from __future__ import annotations

import numpy as np
import pytest
from scipy.sparse import csr_matrix

import oracle
import toy_codes
from pelagos_amd.graph import CSRGraph
from pelagos_amd.utils import compute_log_prior_ratios

DTYPES = ("float64", "float16")
NP = {"float64": np.float64, "float16": np.float16}


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


def _decode(H, p, syn, dtype, gammas=None, **kw):
    g = CSRGraph.from_matrix(csr_matrix(H))
    lam = compute_log_prior_ratios(np.asarray(p, np.float64))  # float32, cast by the oracle
    args = dict(pre_iter=50, num_sets=0, set_max_iter=0, stop_nconv=1, gamma0=0.0, plain=True)
    args.update(kw)
    R = args["num_sets"]
    if gammas is None:
        gammas = np.random.default_rng(9).uniform(-0.24, 0.66, (max(R, 1), g.V)).astype(np.float32)[:R]
    return oracle.decode(g.chk_ptr, g.edge_var, g.V, np.atleast_2d(syn), lam, gammas, dtype=dtype, **args)


def test_half_values_of_every_bit_pattern():
    h = np.arange(1 << 16, dtype=np.uint16).view(np.float16)
    got = oracle.half_value(h)
    exp = h.astype(np.float64)
    nan = np.isnan(exp)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan], exp[~nan])
    assert np.array_equal(np.signbit(got[~nan]), np.signbit(exp[~nan]))


def test_half_rounding_matches_numpy():
    rng = np.random.default_rng(0)
    halves = np.arange(1 << 16, dtype=np.uint16).view(np.float16).astype(np.float64)
    halves = halves[np.isfinite(halves)]
    s = np.sort(np.unique(halves))
    mids = (s[:-1] + s[1:]) / 2                       # ties: round to even
    x = np.concatenate([
        halves, mids, np.nextafter(mids, np.inf), np.nextafter(mids, -np.inf),
        rng.normal(0, 1, 20000) * 10.0 ** rng.integers(-9, 6, 20000),
        [0.0, -0.0, 65504.0, 65519.99, 65520.0, 65536.0, 1e9, -1e9, np.inf, -np.inf,
         2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26, 2.0 ** -26, 2.0 ** -14 - 2.0 ** -25],
    ])
    got = oracle.half_round(x)
    with np.errstate(over="ignore"):
        exp = x.astype(np.float16)
    assert np.array_equal(got.view(np.uint16), exp.view(np.uint16))


@pytest.mark.parametrize("dtype", DTYPES)
def test_first_iteration_closed_form(dtype):
    """Uniform prior, no memory, one iteration: every nu into variable j is
    (-1)^{s_i} lam0 (stored in the element type), and post_j = lam0 + their sum,
    summed from 0 in ascending check order in the sum type (double for float64,
    float for float16), then stored."""
    T = NP[dtype]
    H = toy_codes.regular_code(1, V=60, C=30).toarray()
    lam0 = T(compute_log_prior_ratios(np.array([0.02]))[0])
    rng = np.random.default_rng(2)
    syn = rng.integers(0, 2, (16, 30)).astype(np.uint8)
    out = _decode(H, np.full(60, 0.02), syn, dtype, pre_iter=1, want_posterior=True)
    S = np.float64 if dtype == "float64" else np.float32
    exp = np.zeros((16, 60), T)
    for b in range(16):
        for j in range(60):
            s = S(0.0)
            for i in np.nonzero(H[:, j])[0]:
                nu = T(-lam0 if syn[b, i] else lam0)
                s = S(s + S(nu))
            exp[b, j] = T(S(lam0) + s)
    assert out["last_posterior"].dtype == T
    np.testing.assert_array_equal(out["last_posterior"], exp)


@pytest.mark.parametrize("dtype", DTYPES)
def test_memory_mixing_closed_form(dtype):
    """One variable whose checks send 0: post <- (1 - g) lam0 + g post, each
    operation in the element type (half operations for float16), from post0;
    g is gamma0 as float32, then the element type (decoder.py:3503-3512)."""
    T = NP[dtype]
    H = np.zeros((2, 3), np.uint8)
    H[0, 0] = 1
    H[1, 1] = 1
    H[1, 2] = 1
    g = T(np.float32(0.3))
    lam = compute_log_prior_ratios(np.full(3, 0.05))
    post0 = np.array([[-4.0, 1.5, 2.0]], T)
    gcl = CSRGraph.from_matrix(csr_matrix(H))
    out = oracle.decode(gcl.chk_ptr, gcl.edge_var, 3, np.array([[1, 0]], np.uint8), lam,
                        np.zeros((0, 3), np.float32), pre_iter=3, num_sets=0, set_max_iter=0,
                        stop_nconv=5, gamma0=0.3, init_posterior=post0, want_posterior=True, dtype=dtype)
    p, l0 = T(post0[0, 0]), T(lam[0])
    for _ in range(3):
        # numpy's float16 operations round each result to half (exact for these magnitudes)
        p = T(T(T(T(1.0) - g) * l0) + T(g * p))
    assert out["last_posterior"][0, 0] == p


@pytest.mark.parametrize("dtype", DTYPES)
def test_tree_codes_are_ml(dtype):
    rng = np.random.default_rng(5)
    for n in (5, 8, 11):
        H = toy_codes.repetition_code(n).toarray()
        p = rng.uniform(0.02, 0.3, n)
        w = np.log((1 - p) / p)
        for _ in range(12):
            e = (rng.random(n) < p).astype(np.uint8)
            s = (H @ e) % 2
            out = _decode(H, p, s, dtype, pre_iter=4 * n)
            best, bw = None, np.inf
            for x in range(1 << n):
                c = np.array([(x >> j) & 1 for j in range(n)], np.uint8)
                if np.all((H @ c) % 2 == s) and float((c * w).sum()) < bw - 1e-12:
                    best, bw = c, float((c * w).sum())
            assert out["found_count"][0] == 1
            assert np.array_equal(out["errors"][0], best), (dtype, n, e, s)


@pytest.mark.parametrize("dtype", DTYPES)
def test_relay_solutions_are_valid_with_their_weights(dtype):
    T = NP[dtype]
    rng = np.random.default_rng(6)
    H = toy_codes.random_ldpc(6, 12, 2, seed=1).toarray()
    p = rng.uniform(0.02, 0.15, 12)
    E = (rng.random((32, 12)) < p).astype(np.uint8)
    S = (E @ H.T) % 2
    out = _decode(H, p, S, dtype, pre_iter=20, num_sets=10, set_max_iter=20, stop_nconv=3,
                  gamma0=0.65, plain=False)
    lam = compute_log_prior_ratios(p).astype(T)
    assert out["weights"].dtype == T
    for b in range(S.shape[0]):
        if not np.isfinite(out["weights"][b]):
            continue
        e = out["errors"][b]
        assert np.all((H @ e) % 2 == S[b])
        exact = float((e.astype(np.float64) * lam.astype(np.float64)).sum())
        assert out["weights"][b] == T(np.float32(exact) if dtype == "float16" else exact)
