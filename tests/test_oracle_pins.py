"""Pins for the CPU oracle (oracle/relaybp_oracle.c), independent of the oracle's code.

The relay-bp Rust crate the reference validates against is not available
offline and the reference stores no golden vectors (every expected value in
its tests is computed live from relay_bp or from an invariant).  So each part
of the oracle is pinned by what the mathematics fixes:

* the check update against a brute-force ``prod sign * min(1e30, min over
  the others)`` on exhaustive small rows, ties, zeros and magnitudes beyond
  1e30 included (kernels.py:1120-1215);
* the first iteration against its closed form (uniform priors: every message
  is +-lam0, the posterior is lam0 plus a signed count);
* tree Tanner graphs (chains): min-sum is exact, so the decoded pattern equals
  brute-force maximum likelihood;
* tiny codes: every claimed solution satisfies H e = s, and its weight is at
  least the brute-force ML weight (tests/test_triton_against_rust.py:279-309);
* schedule invariants of the reference tests (iterations of a never-converged
  shot, valid_mask, batch independence, first_leg vs num_sets=0, last_leg);
* the relay-leg schedule (decoder.py:3655-3739, kernels.py:706-741): the
  posterior carried from leg to leg and each leg's own gamma row (a closed
  form across legs), the message reset at every leg start (a gamma-0 relay leg
  replays the plain pre-iter leg bit for bit), and the best-solution rule (a
  lighter later solution replaces the first, an equal-weight one does not);
* planted mistakes: ten mutant builds of the oracle (syndrome sign dropped,
  own edge not excluded, memory weights swapped, extrinsic subtraction
  dropped, messages not reset at a relay leg, posterior reset at a relay leg,
  first solution never replaced, leg l reading row l+1, ties replacing the
  best, check magnitudes not clipped at 1e30) each fail at least one of these
  checks.
"""

from __future__ import annotations

import itertools

import numpy as np
import pytest
from scipy.sparse import csr_matrix

import oracle
import toy_codes
from pelagos_amd import codes
from pelagos_amd.graph import CSRGraph
from pelagos_amd.utils import compute_log_prior_ratios

MUTANTS = {n: oracle._DIR / "_build" / f"liboracle_plant{n}.so" for n in range(1, 11)}


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build(mutants=True)


# --- brute-force references (independent of the oracle) --------------------------------------


CLIP = np.float32(1e30)


def brute_check(mu, syn, alpha=None, beta=None):
    """nu_e = (-1)^s * prod_{k!=e} sgn(mu_k) * min(1e30, min_{k!=e} |mu_k|), sgn(0) = +1
    (the reference's running minima start at 1e30, kernels.py:1120-1126)."""
    mu = np.asarray(mu, np.float32)
    d = mu.shape[0]
    out = np.zeros(d, np.float32)
    for e in range(d):
        others = [k for k in range(d) if k != e]
        if not others:
            out[e] = 0.0
            continue
        sign = -1.0 if syn else 1.0
        for k in others:
            sign *= 1.0 if mu[k] >= 0 else -1.0
        mag = np.float32(min([CLIP] + [abs(mu[k]) for k in others]))
        if alpha is not None:
            mag = np.float32(np.float32(alpha) * mag)
        if beta is not None:
            mag = np.float32(max(mag - np.float32(beta), 0.0))
        out[e] = np.float32(sign) * mag
    return out


def brute_ml(H: np.ndarray, s: np.ndarray, p: np.ndarray):
    """Minimum-weight (max-likelihood) pattern by enumeration (n <= 16)."""
    m, n = H.shape
    w = np.log((1 - p) / p)
    best, best_w = None, np.inf
    for x in range(1 << n):
        e = np.array([(x >> j) & 1 for j in range(n)], np.uint8)
        if np.all((H @ e) % 2 == s):
            ww = float((e * w).sum())
            if ww < best_w - 1e-12:
                best, best_w = e, ww
    return best, best_w


def _decode(H, p, syn, lib_path=None, gammas=None, **kw):
    g = CSRGraph.from_matrix(csr_matrix(H))
    lam = compute_log_prior_ratios(np.asarray(p, np.float64))
    args = dict(pre_iter=50, num_sets=0, set_max_iter=0, stop_nconv=1, gamma0=0.0, plain=True)
    args.update(kw)
    R = args["num_sets"]
    if gammas is None:
        gammas = np.random.default_rng(9).uniform(-0.24, 0.66, (max(R, 1), g.V)).astype(np.float32)[:R]
    return oracle.decode(g.chk_ptr, g.edge_var, g.V, np.atleast_2d(syn), lam, gammas,
                         lib_path=lib_path, **args)


# --- the checks (each takes a library path so mutants can be run through them) -------------


def check_row_brute(lib_path=None):
    rng = np.random.default_rng(0)
    vals = np.array([-2.5, -1.0, -0.0, 0.0, 0.5, 1.0, 2.5], np.float32)
    for d in (1, 2, 3, 4):
        for mu in itertools.product(vals, repeat=d):
            for syn in (0, 1):
                got = oracle.check_update_row(np.array(mu, np.float32), syn, lib_path=lib_path)
                np.testing.assert_array_equal(got, brute_check(mu, syn))
    big = np.array([-3e30, 2e30, 1e30, -1e30, 5e37, np.inf, -np.inf], np.float32)
    for d in (2, 3, 4):
        for mu in itertools.product(np.concatenate([big, vals[[0, 4]]]), repeat=d):
            for syn in (0, 1):
                got = oracle.check_update_row(np.array(mu, np.float32), syn, lib_path=lib_path)
                np.testing.assert_array_equal(got, brute_check(mu, syn))
    for _ in range(200):
        d = int(rng.integers(2, 40))
        mu = rng.normal(0, 3, d).astype(np.float32)
        mu[rng.random(d) < 0.2] = mu[0]           # ties
        syn = int(rng.integers(0, 2))
        a = float(rng.uniform(0.5, 1.0))
        b = float(rng.uniform(0.0, 1.0))
        np.testing.assert_array_equal(oracle.check_update_row(mu, syn, lib_path=lib_path),
                                      brute_check(mu, syn))
        np.testing.assert_array_equal(oracle.check_update_row(mu, syn, alpha=a, lib_path=lib_path),
                                      brute_check(mu, syn, alpha=a))
        np.testing.assert_array_equal(oracle.check_update_row(mu, syn, beta=b, lib_path=lib_path),
                                      brute_check(mu, syn, beta=b))


def check_first_iteration_closed_form(lib_path=None):
    """Uniform prior p, no memory, one iteration, checks of degree >= 2:
    every nu into variable j is (-1)^{s_i} lam0, so post_j = lam0 * (1 + sum_i (-1)^{s_i})
    summed in ascending check order from 0."""
    H = toy_codes.regular_code(1, V=60, C=30).toarray()
    lam0 = compute_log_prior_ratios(np.array([0.02]))[0]
    rng = np.random.default_rng(2)
    syn = rng.integers(0, 2, (16, 30)).astype(np.uint8)
    out = _decode(H, np.full(60, 0.02), syn, lib_path, pre_iter=1, want_posterior=True)
    exp = np.zeros((16, 60), np.float32)
    for b in range(16):
        for j in range(60):
            s = np.float32(0.0)
            for i in np.nonzero(H[:, j])[0]:
                s = np.float32(s + (-lam0 if syn[b, i] else lam0))
            exp[b, j] = np.float32(lam0 + s)
    np.testing.assert_array_equal(out["last_posterior"], exp)


def check_second_iteration_closed_form(lib_path=None):
    """Triangle code [[1,1,0],[0,1,1]], uniform LLR l, s = (1, 1), no memory.

    Iteration 1: every check sends -l, so post = (0, -l, 0) (hard 1,1,1 fails
    check 0).  mu = lam + (S - nu) gives mu(c0,v0) = l, mu(c0,v1) = 0,
    mu(c1,v1) = 0, mu(c1,v2) = l; iteration 2 sends (-0, -l) on c0 and (-l, -0)
    on c1, so post = (l, -l, l): the single error on v1, found at iteration 2.
    """
    H = np.array([[1, 1, 0], [0, 1, 1]], np.uint8)
    l = compute_log_prior_ratios(np.array([0.1]))[0]
    s = np.array([1, 1], np.uint8)
    o1 = _decode(H, np.full(3, 0.1), s, lib_path, pre_iter=1, want_posterior=True)
    o2 = _decode(H, np.full(3, 0.1), s, lib_path, pre_iter=2, want_posterior=True)
    np.testing.assert_array_equal(o1["last_posterior"][0], np.array([0.0, -l, 0.0], np.float32))
    assert o1["found_count"][0] == 0 and o1["iterations"][0] == 1
    np.testing.assert_array_equal(o2["last_posterior"][0], np.array([l, -l, l], np.float32))
    assert o2["found_count"][0] == 1 and o2["iterations"][0] == 2
    np.testing.assert_array_equal(o2["errors"][0], np.array([0, 1, 0], np.uint8))
    assert o2["weights"][0] == l


def check_memory_mixing_closed_form(lib_path=None):
    """One isolated variable with two degree-1 checks (which send 0): the
    posterior follows post <- (1-g)*lam0 + g*post, two roundings, from post0."""
    H = np.zeros((2, 3), np.uint8)
    H[0, 0] = 1
    H[1, 1] = 1
    H[1, 2] = 1
    g = np.float32(0.3)
    lam = compute_log_prior_ratios(np.full(3, 0.05))
    post0 = np.array([[-4.0, 1.5, 2.0]], np.float32)
    g0 = None
    gcl = CSRGraph.from_matrix(csr_matrix(H))
    out = oracle.decode(gcl.chk_ptr, gcl.edge_var, 3, np.array([[1, 0]], np.uint8), lam,
                        np.zeros((0, 3), np.float32), pre_iter=3, num_sets=0, set_max_iter=0,
                        stop_nconv=5, gamma0=float(g), init_posterior=post0, want_posterior=True,
                        lib_path=lib_path)
    p = np.float32(post0[0, 0])
    for _ in range(3):
        p = np.float32(np.float32(np.float32(1.0) - g) * lam[0]) + np.float32(g * p)
        p = np.float32(p + np.float32(0.0))
    assert out["last_posterior"][0, 0] == p
    del g0


def check_tree_codes_are_ml(lib_path=None):
    """Chain codes (tree Tanner graphs): plain min-sum returns the ML pattern."""
    rng = np.random.default_rng(5)
    for n in (5, 8, 11):
        H = toy_codes.repetition_code(n).toarray()
        p = rng.uniform(0.02, 0.3, n)
        for _ in range(12):
            e = (rng.random(n) < p).astype(np.uint8)
            s = (H @ e) % 2
            out = _decode(H, p, s, lib_path, pre_iter=4 * n)
            ml, wml = brute_ml(H, s, p)
            assert out["found_count"][0] == 1
            assert np.array_equal(out["errors"][0], ml), (n, e, s)


def check_tiny_codes_valid_and_not_below_ml(lib_path=None):
    rng = np.random.default_rng(6)
    for trial in range(6):
        n, m = 12, 6
        H = toy_codes.random_ldpc(m, n, 2, seed=trial).toarray()
        p = rng.uniform(0.02, 0.15, n)
        E = (rng.random((24, n)) < p).astype(np.uint8)
        S = (E @ H.T) % 2
        out = _decode(H, p, S, lib_path, pre_iter=20, num_sets=10, set_max_iter=20, stop_nconv=3,
                      gamma0=0.65, plain=False)
        lam = compute_log_prior_ratios(p)
        for b in range(S.shape[0]):
            if not np.isfinite(out["weights"][b]):
                continue
            e = out["errors"][b]
            assert np.all((H @ e) % 2 == S[b])
            _, wml = brute_ml(H, S[b], p)
            assert float((e * lam).sum()) >= wml - 1e-3
            assert abs(float(out["weights"][b]) - float((e.astype(np.float64) * lam).sum())) < 1e-4


def check_schedule_invariants(lib_path=None):
    H = toy_codes.regular_code(3, V=120, C=60).toarray()
    p = np.full(120, 0.03)
    rng = np.random.default_rng(7)
    E = (rng.random((40, 120)) < p * 3).astype(np.uint8)
    S = (E @ H.T) % 2
    kw = dict(pre_iter=10, num_sets=6, set_max_iter=8, stop_nconv=3, gamma0=0.65, plain=False)
    full = _decode(H, p, S, lib_path, **kw)
    # valid_mask == found >= stop_nconv; solved shots satisfy H e = s
    assert np.array_equal(full["valid_mask"], full["found_count"] >= 3)
    solved = np.isfinite(full["weights"])
    assert np.all(((full["errors"][solved] @ H.T) % 2) == S[solved])
    # batch independence
    sub = _decode(H, p, S[5:9], lib_path, **kw)
    for k in ("errors", "iterations", "found_count", "weights"):
        assert np.array_equal(full[k][5:9], sub[k])
    # iterations never exceed the schedule; a shot with found 0 ran every iteration
    assert full["iterations"].max() <= 10 + 6 * 8
    # unsatisfiable syndrome (an empty check with s = 1)
    H2 = np.zeros((3, 6), np.uint8)
    H2[0, :3] = 1
    H2[1, 3:] = 1
    o = _decode(H2, np.full(6, 0.05), np.array([[1, 0, 1]], np.uint8), lib_path, **kw)
    assert o["iterations"][0] == 10 + 6 * 8 and o["found_count"][0] == 0
    # first_leg correction = the num_sets = 0 decode's hard decision
    kw0 = dict(kw, num_sets=0)
    o0 = _decode(H, p, S, lib_path, **kw0)
    never = ~np.isfinite(full["weights"])
    assert np.array_equal(full["errors"][never], o0["errors"][never])


def _mix(g, lam0, post):
    """lam_mixed = (1-g)*lam0 + g*post with the kernel's two roundings (kernels.py:1309)."""
    g = np.float32(g)
    return np.float32(np.float32(np.float32(np.float32(1.0) - g) * np.float32(lam0)) + np.float32(g * np.float32(post)))


def check_relay_legs_carry_the_posterior_with_their_own_rows(lib_path=None):
    """Check {v0, v1} with syndrome 1 and an empty check with syndrome 1 (never
    satisfied, so every leg runs).  v1 has degree 1 and a zero strength in
    every leg, so its message is its prior l1 and v0 receives the constant
    c = -l1: v0 follows post <- (1-g)*l0 + g*post + c, whose fixed point
    depends on g, so the legs' order matters.  The pre-iter leg uses init_gamma
    (0.3 for v0), then relay leg l uses row l of the table, each leg starting
    from the posterior the previous one left (decoder.py:3672-3700).  A reset
    posterior or a shifted row changes the final value."""
    H = np.zeros((2, 2), np.uint8)
    H[0, :] = 1
    gcl = CSRGraph.from_matrix(csr_matrix(H))
    lam = np.array([2.0, 0.75], np.float32)
    rows = np.array([[0.5, 0.0], [-0.2, 0.0], [0.1, 0.0], [0.62, 0.0]], np.float32)
    post0 = np.array([[-4.0, 0.75]], np.float32)
    T0, Tr = 3, 2
    out = oracle.decode(gcl.chk_ptr, gcl.edge_var, 2, np.array([[1, 1]], np.uint8), lam, rows,
                        pre_iter=T0, num_sets=rows.shape[0], set_max_iter=Tr, stop_nconv=1, gamma0=0.65,
                        init_gamma=np.array([[0.3, 0.0]], np.float32), init_posterior=post0,
                        want_posterior=True, lib_path=lib_path)
    c = np.float32(np.float32(0.0) + np.float32(-0.75))
    p = np.float32(post0[0, 0])
    for _ in range(T0):
        p = np.float32(_mix(0.3, lam[0], p) + c)
    for leg in range(rows.shape[0]):
        for _ in range(Tr):
            p = np.float32(_mix(rows[leg, 0], lam[0], p) + c)
    assert out["found_count"][0] == 0 and out["iterations"][0] == T0 + rows.shape[0] * Tr
    assert out["last_posterior"][0, 0] == p


def _never_converging_case():
    """A loopy (3,6)-regular code plus one empty check whose syndrome bit is
    1, so no error pattern explains the syndrome and BP runs every
    iteration."""
    H0 = toy_codes.regular_code(3, V=48, C=24).toarray()
    H = np.vstack([H0, np.zeros((1, H0.shape[1]), np.uint8)])
    rng = np.random.default_rng(12)
    e = (rng.random((8, H.shape[1])) < 0.15).astype(np.uint8)
    S = (e @ H.T) % 2
    S[:, -1] = 1
    return H, np.full(H.shape[1], 0.04), S


def check_relay_leg_resets_the_messages(lib_path=None):
    """With a gamma-0 row a relay leg mixes nothing of the posterior in, so,
    because every leg starts from messages reset to the priors
    (init_messages_with_priors_EB_kernel, kernels.py:903-1000; decoder.py:
    3729-3739), it replays the plain pre-iter leg bit for bit.  Continuing the
    pre-iter leg's messages instead gives the 2T-iteration trajectory, which
    differs."""
    H, p, S = _never_converging_case()
    T = 9
    zero_row = np.zeros((1, H.shape[1]), np.float32)
    a = _decode(H, p, S, lib_path, pre_iter=T, want_posterior=True)
    b = _decode(H, p, S, lib_path, pre_iter=T, num_sets=1, set_max_iter=T, gammas=zero_row, want_posterior=True)
    c = _decode(H, p, S, lib_path, pre_iter=2 * T, want_posterior=True)
    assert (a["found_count"] == 0).all() and (b["iterations"] == 2 * T).all()
    assert not np.array_equal(a["last_posterior"], c["last_posterior"]), "trajectory must move"
    np.testing.assert_array_equal(b["last_posterior"], a["last_posterior"])


def check_lighter_solution_replaces_the_first_and_ties_do_not(lib_path=None):
    """Two checks {v0,v1}, {v2,v3} with syndrome (1,1) and LLRs (1,2,2,1):
    four solutions of weights {0,2}=3, {0,3}=2, {1,2}=4, {1,3}=3.  The
    pre-iter leg (gamma0 = 1, so the first iteration's mix is the initial
    posterior) is driven to {0,2}.  A relay leg with a zero row then converges
    at once to the lighter {0,3}, which replaces it; a relay leg whose row
    (-0.2, -0.5, 0, 0) mixes in the carried posterior converges to {1,3}, whose
    equal weight keeps {0,2} (first solution, then strictly lower weight:
    kernels.py:706-741)."""
    H = np.array([[1, 1, 0, 0], [0, 0, 1, 1]], np.uint8)
    gcl = CSRGraph.from_matrix(csr_matrix(H))
    lam = np.array([1.0, 2.0, 2.0, 1.0], np.float32)
    syn = np.array([[1, 1]], np.uint8)
    post0 = np.array([[-10.0, 10.0, -10.0, 10.0]], np.float32)
    kw = dict(pre_iter=5, num_sets=1, set_max_iter=5, stop_nconv=2, gamma0=1.0, init_posterior=post0,
              lib_path=lib_path)
    first = oracle.decode(gcl.chk_ptr, gcl.edge_var, 4, syn, lam, np.zeros((0, 4), np.float32),
                          **dict(kw, num_sets=0, stop_nconv=1))
    assert np.array_equal(first["errors"][0], [1, 0, 1, 0]) and first["weights"][0] == 3.0
    lighter = oracle.decode(gcl.chk_ptr, gcl.edge_var, 4, syn, lam, np.zeros((1, 4), np.float32), **kw)
    assert lighter["found_count"][0] == 2 and lighter["iterations"][0] == 2
    assert np.array_equal(lighter["errors"][0], [1, 0, 0, 1]) and lighter["weights"][0] == 2.0
    tie = oracle.decode(gcl.chk_ptr, gcl.edge_var, 4, syn, lam, np.array([[-0.2, -0.5, 0.0, 0.0]], np.float32),
                        want_posterior=True, **kw)
    assert tie["found_count"][0] == 2 and tie["iterations"][0] == 2
    assert np.array_equal((tie["last_posterior"][0] <= 0).astype(np.uint8), [0, 1, 0, 1])  # found {1,3}
    assert np.array_equal(tie["errors"][0], [1, 0, 1, 0]) and tie["weights"][0] == 3.0


def check_pre_iter_zero_falls_back_to_relay_leg_one(lib_path=None):
    """pre_iter = 0: a shot without a solution returns the hard decision after
    relay leg 1, which is what a one-leg last_leg decode returns
    (tests/test_unconverged_correction.py:124-139)."""
    H, p, S = _never_converging_case()
    rows = np.random.default_rng(3).uniform(-0.24, 0.66, (3, H.shape[1])).astype(np.float32)
    full = _decode(H, p, S, lib_path, pre_iter=0, num_sets=3, set_max_iter=7, gammas=rows,
                   gamma0=0.65, plain=False)
    leg1 = _decode(H, p, S, lib_path, pre_iter=0, num_sets=1, set_max_iter=7, gammas=rows[:1],
                   gamma0=0.65, plain=False, last_leg=True)
    assert (full["found_count"] == 0).all()
    np.testing.assert_array_equal(full["errors"], leg1["errors"])
    assert full["errors"].any()


ALL_CHECKS = [
    check_row_brute,
    check_first_iteration_closed_form,
    check_second_iteration_closed_form,
    check_memory_mixing_closed_form,
    check_tree_codes_are_ml,
    check_tiny_codes_valid_and_not_below_ml,
    check_schedule_invariants,
    check_relay_legs_carry_the_posterior_with_their_own_rows,
    check_relay_leg_resets_the_messages,
    check_lighter_solution_replaces_the_first_and_ties_do_not,
    check_pre_iter_zero_falls_back_to_relay_leg_one,
]


@pytest.mark.parametrize("check", ALL_CHECKS, ids=lambda c: c.__name__)
def test_oracle_pinned(check):
    check()


@pytest.mark.parametrize("plant", sorted(MUTANTS))
def test_planted_mistake_is_caught(plant):
    """Each mutant oracle fails at least one pinning check."""
    lib = MUTANTS[plant]
    assert oracle.load(lib).orc_planted() == plant
    failures = 0
    for check in ALL_CHECKS:
        try:
            check(lib)
        except AssertionError:
            failures += 1
    assert failures > 0, f"planted mistake {plant} passed every pinning check"
