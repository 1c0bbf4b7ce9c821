"""Learned memory strengths: the training kernels against the CPU restatement
(oracle/mlp_oracle.py, itself pinned by finite differences), the matrix-core
inference inside the decode kernel against the same network in bf16, and the
learned decode against the CPU oracle fed the strengths it used."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from oracle import mlp_oracle as mo
from pelagos_amd import codes, learned

P = mo.P


def _problem(n=3000, seed=0):
    rng = np.random.default_rng(seed)
    p = learned.init_params(seed)
    p[:] += rng.normal(0, 0.05, P).astype(np.float32)  # non-trivial biases
    x = rng.normal(0, 1, (n, 8)).astype(np.float32)
    y = (rng.random(n) < 0.7).astype(np.float32)
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    return p, x, y, w


# --- CPU: pins of the restatement ------------------------------------------------------------


def test_parameter_count():
    assert P == 32 * 8 + 32 + 32 * 32 + 32 + 32 + 1 == learned.N_PARAMS


def test_oracle_gradient_matches_finite_differences():
    p, x, y, w = _problem(200, seed=1)
    g, _ = mo.grad(p.astype(np.float64), x, y, w)
    rng = np.random.default_rng(2)
    for k in rng.choice(P, 60, replace=False).tolist() + [P - 1, 0, 300]:
        e = np.zeros(P)
        e[k] = 1e-5
        num = (mo.loss(p + e, x, y, w) - mo.loss(p - e, x, y, w)) / 2e-5
        assert abs(num - g[k]) <= 1e-6 + 1e-4 * abs(num), (k, num, g[k])


def test_oracle_loss_closed_form_at_zero_network():
    """All weights zero: z = b3 everywhere, loss = mean w (softplus(b3) - y b3)."""
    p = np.zeros(P)
    p[-1] = 0.3
    _, x, y, w = _problem(50, seed=3)
    y, w = y.astype(np.float64), w.astype(np.float64)
    exp = np.mean(w * (np.log1p(np.exp(0.3)) - y * 0.3))
    assert abs(mo.loss(p, x, y, w) - exp) < 1e-12
    g, _ = mo.grad(p, x, y, w)
    assert abs(g[-1] - np.mean(w * (1 / (1 + np.exp(-0.3)) - y))) < 1e-12
    assert not g[:-1].any()


def _bf16_exact_network(seed):
    """Features in multiples of 1/4 on [-1, 1] (feature 7, the leg index, 0 as
    at relay leg 0), W1 entries in {-1, 0, 1}/4, W2 rows with at most four
    entries in {-1, 1}/8, w3 in multiples of 1/8, dyadic biases: every
    activation has at most 8 significant bits, so the kernel's bf16 roundings
    (features, weights, both layers' activations) are all exact and its
    inference must equal the float64 network."""
    rng = np.random.default_rng(seed)
    W1 = rng.integers(-1, 2, (32, 8)) / 4.0
    b1 = rng.integers(-4, 5, 32) / 8.0
    W2 = np.zeros((32, 32))
    for r in range(32):
        W2[r, rng.choice(32, 4, replace=False)] = rng.choice([-1.0, 1.0], 4) / 8.0
    b2 = rng.integers(-4, 5, 32) / 16.0
    w3 = rng.integers(-8, 9, 32) / 8.0
    p = np.concatenate([W1.ravel(), b1, W2.ravel(), b2, w3, [0.1]]).astype(np.float32)
    x = (rng.integers(-4, 5, (4000, 8)) / 4.0).astype(np.float32)
    x[:, 7] = 0.0
    return p, x


def test_oracle_bf16_inference_matches_the_network_when_rounding_is_exact():
    """shift_q8 (the decode kernel's bf16 inference, rounded to int8 steps of
    (hi - lo) / 254) against forward() in float64: on a network and features
    whose bf16 roundings are exact, the steps agree wherever 254 (sigmoid(z) -
    1/2) is not within 1e-4 of a rounding tie.  A transposed W2, swapped
    biases or a dropped relu change the logits of this network."""
    lo, hi = -0.24, 0.66
    for seed in (0, 1, 2):
        p, x = _bf16_exact_network(seed)
        assert np.array_equal(mo.bf16(p[:32 * 8]), p[:32 * 8]) and np.array_equal(mo.bf16(x), x)
        z, c = mo.forward(p.astype(np.float64), x)
        for k in ("h1", "h2"):
            assert np.array_equal(mo.bf16(c[k].astype(np.float32)), c[k].astype(np.float32)), k
        t = 254.0 * (1.0 / (1.0 + np.exp(-z)) - 0.5)
        steps_ref = np.rint(t)
        steps = np.rint(mo.shift_q8(p, x, lo, hi) / np.float32((hi - lo) / 254.0))
        clear = np.abs(np.abs(t - np.floor(t)) - 0.5) > 1e-4
        assert clear.mean() > 0.99 and np.abs(steps_ref).max() > 20
        np.testing.assert_array_equal(steps[clear], steps_ref[clear])
        assert np.abs(steps - steps_ref).max() <= 1
        # a plausible slip: W2 transposed gives other steps on this network
        q = p.copy()
        q[32 * 8 + 32:32 * 8 + 32 + 1024] = p[32 * 8 + 32:32 * 8 + 32 + 1024].reshape(32, 32).T.ravel()
        assert (np.rint(mo.shift_q8(q, x, lo, hi) / np.float32((hi - lo) / 254.0)) != steps).mean() > 0.1
        # and so does a dropped first-layer bias (the kernel carries it in feature 7's slot)
        q = p.copy()
        q[32 * 8:32 * 8 + 32] = 0.0
        assert (np.rint(mo.shift_q8(q, x, lo, hi) / np.float32((hi - lo) / 254.0)) != steps).mean() > 0.1


def _rn32(num: int, exp: int):
    """num * 2**exp rounded to float32 (nearest, ties to even) as (mantissa, exponent)
    with a mantissa of at most 24 bits; exact integer arithmetic, normal range."""
    if num == 0:
        return 0, 0
    sign, n = (-1 if num < 0 else 1), abs(num)
    shift = n.bit_length() - 24
    if shift > 0:
        q, rem, half = n >> shift, n & ((1 << shift) - 1), 1 << (shift - 1)
        if rem > half or (rem == half and q & 1):
            q += 1
        n, exp = q, exp + shift
    return sign * n, exp


def test_reciprocal_with_one_newton_step_is_correctly_rounded():
    """The feature pass takes 1/deg as r1 = fma(r0, fma(-deg, r0, 1), r0) from the
    hardware reciprocal r0 (within 1 ulp of 1/deg).  For every integer deg in
    [1, 65535] and every float within 1 ulp of 1/deg (and the float just below
    a power of two), r1 is 1/deg rounded to the nearest float32: the features
    equal those of a correctly rounded division."""
    def eq(a, b):  # (mantissa, exponent) pairs as values
        (ma, ea), (mb, eb) = a, b
        e = min(ea, eb)
        return ma << (ea - e) == mb << (eb - e)

    for d in range(1, 65536):
        k = 23 + d.bit_length()
        q, r = divmod(1 << k, d)  # 2**k / d in (2**23, 2**24]
        if 2 * r > d or (2 * r == d and q & 1):
            q += 1
        if q == 1 << 24:
            q, k = 1 << 23, k - 1
        rn = (q, -k)
        cands = [(q - 1, -k), (q, -k), (q + 1, -k)]
        if q == 1 << 23:
            cands.append(((1 << 24) - 1, -k - 1))
        for m0, e0 in cands:
            # e = RN(1 - d r0): 1 - d m0 2**e0 = ((1 << -e0) - d m0) 2**e0, exact
            me, ee = _rn32((1 << -e0) - d * m0, e0)
            # r1 = RN(r0 + r0 e) = m0 (2**-ee + me) 2**(e0 + ee), with ee < 0 (or e = 0)
            assert me == 0 or ee < 0
            r1 = (m0, e0) if me == 0 else _rn32(m0 * ((1 << -ee) + me), e0 + ee)
            assert eq(r1, rn), (d, m0, e0)


def test_oracle_leg0_features_against_the_decode_state():
    """The oracle's leg-0 features against what its pinned decode fixes: prior
    and degree closed forms, the posterior the pre-iter leg ends with, no flips
    beyond its iterations, violated-check fractions recomputed from that
    posterior's hard decisions, and no features for shots that stop before
    relay leg 0."""
    H, p = codes.banded_circuit_code(n_blocks=4, V=2700, seed=2)
    err = codes.sample_errors(p, 48, seed=12, scale=2.0)
    syn = codes.syndromes_of(H, err)
    Hc = H.tocsr()
    V = H.shape[1]
    lam = np.log((1 - p) / p).astype(np.float32)
    table = np.random.default_rng(0).uniform(-0.24, 0.66, (12, V)).astype(np.float32)
    kw = dict(pre_iter=30, num_sets=12, set_max_iter=20, stop_nconv=4, gamma0=0.65)
    f = oracle.decode(Hc.indptr, Hc.indices, V, syn, lam, table, want_features0=True, **kw)["features0"]
    pre = oracle.decode(Hc.indptr, Hc.indices, V, syn, lam, table, skip_relay_legs=True, want_posterior=True, **kw)
    r = pre["found_count"] < 4
    conv = pre["found_count"] >= 1
    assert r.sum() > 10 and (r & ~conv).sum() > 0 and (r & conv).sum() > 0
    assert not f[~r].any()
    deg = np.diff(Hc.tocsc().indptr).astype(np.float32)
    np.testing.assert_array_equal(f[r][..., 0], np.broadcast_to(lam * np.float32(0.125), (r.sum(), V)))
    np.testing.assert_array_equal(f[r][..., 5], np.broadcast_to(deg * np.float32(0.125), (r.sum(), V)))
    assert not f[r][..., 7].any()
    post = pre["last_posterior"]
    np.testing.assert_array_equal(f[r][..., 1], np.clip(post[r] * np.float32(0.0625), -4, 4))
    flips = f[..., 2] * 8
    assert np.array_equal(flips, np.round(flips)) and (flips <= pre["iterations"][:, None]).all() and flips.max() > 0
    # violated checks from the posterior's hard decisions (checks x variables incidence)
    hard = (post <= 0).astype(np.int64)
    viol = ((hard @ Hc.T.toarray().astype(np.int64)) + syn) % 2         # [B, C]
    frac = (viol @ Hc.toarray().astype(np.int64)) / np.maximum(deg, 1)   # [B, V]
    np.testing.assert_allclose(f[r & ~conv][..., 6], frac[r & ~conv], rtol=1e-6)
    assert not f[r & conv][..., 6].any() and f[r & ~conv][..., 6].max() > 0
    # the smallest incoming message never exceeds the clamp
    assert (np.abs(f[r][..., 4]) <= 4).all()
    # features 3 and 4 from the messages the pre-iter leg ends with: the sum of the last variable update's
    # incoming messages in ascending edge order, and the smallest message a next variable update would read
    # (after a leg that converged before its last iteration, one more check pass over the final mu, by the
    # separately pinned single-check update; otherwise the last check pass's)
    msg = oracle.decode(Hc.indptr, Hc.indices, V, syn, lam, table, want_messages0=True, **kw)
    order = np.argsort(Hc.indices, kind="stable")               # edges by variable, ascending edge id
    starts = np.concatenate([[0], np.cumsum(deg.astype(np.int64))])
    inv = np.where(deg > 0, np.float32(1) / np.maximum(deg, 1), np.float32(0)).astype(np.float32)
    C = Hc.shape[0]
    early = conv & (pre["iterations"] < kw["pre_iter"])
    assert (r & early).sum() > 5
    for b in np.flatnonzero(r)[:16]:
        nu_last = msg["nu0"][b]
        nu_next = nu_last
        if early[b]:
            nu_next = np.concatenate([oracle.check_update_row(msg["mu0"][b, Hc.indptr[i]:Hc.indptr[i + 1]], syn[b, i])
                                      for i in range(C)])
        s_sum = np.zeros(V, np.float32)
        mn = np.full(V, np.inf, np.float32)
        for k in range(int(deg.max())):
            has = deg > k
            e = order[np.minimum(starts[:-1] + k, len(order) - 1)]
            s_sum = np.where(has, s_sum + np.where(has, nu_last[e], 0), s_sum).astype(np.float32)
            mn = np.where(has, np.minimum(mn, nu_next[e]), mn).astype(np.float32)
        np.testing.assert_array_equal(f[b, :, 3], np.clip(s_sum * inv * np.float32(0.125), -4, 4))
        np.testing.assert_array_equal(f[b, :, 4], np.where(deg > 0, np.clip(mn * np.float32(0.125), -4, 4), 0))
    # with a plain pre-iter leg (strength 0) the memory term is the prior itself, so the incoming
    # sum of the last variable update is post - prior (up to the rounding of post)
    kw0 = dict(kw, plain=True)
    f0 = oracle.decode(Hc.indptr, Hc.indices, V, syn, lam, table, want_features0=True, **kw0)["features0"]
    pre0 = oracle.decode(Hc.indptr, Hc.indices, V, syn, lam, table, skip_relay_legs=True, want_posterior=True, **kw0)
    r0 = pre0["found_count"] < 4
    post0 = pre0["last_posterior"].astype(np.float64)
    s_exp = np.clip((post0 - lam[None, :]) / deg / 8, -4, 4)
    tol = 1e-6 * np.maximum(np.abs(lam)[None, :], np.abs(post0)) / deg + 1e-7
    assert r0.sum() > 10
    assert (np.abs(f0[..., 3] - s_exp) <= tol)[r0].all()


# --- GPU: training kernels ------------------------------------------------------------------


@pytest.mark.gpu
def test_gpu_gradient_matches_oracle_and_is_reproducible():
    p, x, y, w = _problem(5000, seed=4)
    net = learned.GammaMLP(p)
    X, Y, W = (torch.from_numpy(a).cuda() for a in (x, y, w))
    print("[gpu] mlp grad n=5000", flush=True)
    g1, l1 = net.grad(X, Y, W)
    g2, l2 = net.grad(X, Y, W)
    torch.cuda.synchronize()
    assert torch.equal(g1, g2) and torch.equal(l1, l2)
    ge, le = mo.grad(p, x, y, w)
    np.testing.assert_allclose(g1.cpu().numpy(), ge, rtol=2e-4, atol=2e-6)
    assert abs(float(l1[0]) - le) <= 1e-5 * abs(le) + 1e-6


@pytest.mark.gpu
def test_gpu_forward_and_adam_match_the_formulas():
    p, x, y, w = _problem(4000, seed=5)
    net = learned.GammaMLP(p)
    z = net.logits(torch.from_numpy(x).cuda()).cpu().numpy()
    ze, _ = mo.forward(p, x)
    np.testing.assert_allclose(z, ze, rtol=1e-4, atol=1e-5)
    g = torch.from_numpy(np.random.default_rng(6).normal(0, 1, P).astype(np.float32)).cuda()
    net.adam_step(g, lr=0.01)
    torch.cuda.synchronize()
    gn = g.cpu().numpy().astype(np.float64)
    m1, m2 = 0.1 * gn, 0.001 * gn * gn
    exp = p - 0.01 * (m1 / 0.1) / (np.sqrt(m2 / 0.001) + 1e-8)
    np.testing.assert_allclose(net.params.cpu().numpy(), exp, rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
def test_gpu_training_reduces_the_loss():
    rng = np.random.default_rng(7)
    x = rng.normal(0, 1, (20000, 8)).astype(np.float32)
    y = ((x[:, 0] - 0.5 * x[:, 3] + 0.2 * rng.normal(size=20000)) > 0).astype(np.float32)
    net = learned.GammaMLP(seed=1)
    X, Y, W = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), torch.ones(20000, device="cuda")
    print("[gpu] mlp fit", flush=True)
    losses = net.fit(X, Y, W, epochs=6, batch=2048, lr=5e-3)
    assert losses[-1] < 0.5 * losses[0], losses


# --- GPU: the learned decode ------------------------------------------------------------------


def _banded(shots, seed):
    H, p = codes.banded_circuit_code(n_blocks=4, V=2700, seed=2)
    err = codes.sample_errors(p, shots, seed=seed, scale=2.0)
    return H, p, err, codes.syndromes_of(H, err)


def _decoder(H, p, **kw):
    from pelagos_amd import RelayBPDecoder

    return RelayBPDecoder(H, p, pre_iter=30, num_sets=12, set_max_iter=20, gamma0=0.65,
                          gamma_dist_interval=(-0.24, 0.66), stop_nconv=4, data_type="float32", algo="relay",
                          seed=3, device=[0], **kw)


@pytest.mark.gpu
@pytest.mark.parametrize("requeue,placement", [(0, None), (None, None), (None, 7)])
def test_gpu_learned_decode_matches_oracle_given_its_strengths(requeue, placement):
    """The learned decode equals the CPU oracle run with the strengths the
    network gave (per shot, per learned leg) and the uniform rows after them,
    with and without re-queueing, with the priors in LDS (placement 15, the
    default here) and read from L2 (placement 7)."""
    H, p, err, syn = _banded(96, seed=11)
    L = 3
    kw = {} if placement is None else {"state_placement": placement}
    dec = _decoder(H, p, learned_gamma=learned.GammaMLP(seed=2), learned_legs=L, **kw)
    if requeue is not None:
        dec.requeue_below = requeue
    dec.record_learned_gammas = True
    print(f"[gpu] learned decode requeue={requeue} placement={placement}", flush=True)
    out = dec.decode(torch.from_numpy(syn).cuda())
    torch.cuda.synchronize()
    gl = dec.learned_gammas.cpu().numpy()                      # [B, L, V]
    B, V = syn.shape[0], H.shape[1]
    rows = dec.gamma_table(B).cpu().numpy()                    # [R, V]
    table = np.broadcast_to(rows[:, None, :], (rows.shape[0], B, V)).copy()
    table[:L] = np.transpose(gl, (1, 0, 2))
    lam = dec.log_prior_ratios_per_gpu[0].reshape(-1).cpu().numpy()
    ref = oracle.decode(dec.graph.chk_ptr, dec.graph.edge_var, V, syn, lam, table, pre_iter=30, num_sets=12,
                        set_max_iter=20, stop_nconv=4, gamma0=0.65, gamma_per_row=True)
    for k in ("errors", "iterations", "found_count", "weights"):
        np.testing.assert_array_equal(out[k].cpu().numpy(), ref[k], err_msg=k)
    # every learned strength lies in the interval
    used = gl[gl != 0]
    assert used.size and used.min() >= np.float32(-0.24) and used.max() <= np.float32(0.66)


@pytest.mark.gpu
def test_gpu_last_learned_leg_damping():
    """learned_last_damp scales the shift of the last learned leg only: with
    damp 1 a fourth learned leg is a uniform leg (the decode equals three
    learned legs bit for bit); with damp 1/2 the decode equals the oracle fed
    the strengths it recorded, and leg 3's unclamped strengths sit at half of
    leg 0's shift from their own draws."""
    from pelagos_amd import RelayBPDecoder

    def make(**kw):  # five solutions: the pre-iter leg and four relay legs
        return RelayBPDecoder(H, p, pre_iter=30, num_sets=12, set_max_iter=20, gamma0=0.65,
                              gamma_dist_interval=(-0.24, 0.66), stop_nconv=5, data_type="float32",
                              algo="relay", seed=3, device=[0], **kw)

    H, p, err, syn = _banded(96, seed=15)
    S = torch.from_numpy(syn).cuda()
    net = learned.GammaMLP(seed=2)
    print("[gpu] last-leg damping", flush=True)
    three = make(learned_gamma=net, learned_legs=3).decode(S)
    full_damp = make(learned_gamma=net, learned_legs=4, learned_last_damp=1.0).decode(S)
    for k in ("errors", "iterations", "found_count", "weights"):
        assert torch.equal(three[k], full_damp[k]), k
    L = 4
    dec = make(learned_gamma=net, learned_legs=L, learned_last_damp=0.5)
    dec.record_learned_gammas = True
    out = dec.decode(S)
    torch.cuda.synchronize()
    gl = dec.learned_gammas.cpu().numpy()                      # [B, L, V]
    B, V = syn.shape[0], H.shape[1]
    rows = dec.gamma_table(B).cpu().numpy()                    # [R, V]
    table = np.broadcast_to(rows[:, None, :], (rows.shape[0], B, V)).copy()
    table[:L] = np.transpose(gl, (1, 0, 2))
    lam = dec.log_prior_ratios_per_gpu[0].reshape(-1).cpu().numpy()
    ref = oracle.decode(dec.graph.chk_ptr, dec.graph.edge_var, V, syn, lam, table, pre_iter=30, num_sets=12,
                        set_max_iter=20, stop_nconv=5, gamma0=0.65, gamma_per_row=True)
    for k in ("errors", "iterations", "found_count", "weights"):
        np.testing.assert_array_equal(out[k].cpu().numpy(), ref[k], err_msg=k)
    lo, hi = np.float32(-0.24), np.float32(0.66)
    d0 = gl[:, 0] - rows[0][None, :]
    d3 = gl[:, 3] - rows[3][None, :]
    ran = (gl[:, 3] != 0).any(1)                # rows of legs a shot never reached stay 0
    reached = ran & (np.abs(d3) > 0).any(1)      # shots whose leg 3 ran shifted (so did their leg 0)
    inner = reached[:, None] & (gl[:, 0] > lo) & (gl[:, 0] < hi) & (gl[:, 3] > lo) & (gl[:, 3] < hi)
    assert inner.sum() > 100
    np.testing.assert_allclose(d3[inner], 0.5 * d0[inner], atol=2e-6)


@pytest.mark.gpu
def test_gpu_inference_matches_the_network_in_bf16():
    """Strengths of relay leg 0 (features from the pre-iter leg, the same in
    the plain and the learned decode) against the restatement with bf16
    operands: clamp(u + shift(x), lo, hi) with u the leg's uniform draw, for
    shots whose first leg converged; the others run leg 0 on u itself."""
    H, p, err, syn = _banded(64, seed=12)
    net = learned.GammaMLP(seed=4)
    plain = _decoder(H, p)
    tr = plain.collect_traces(torch.from_numpy(syn).cuda(), legs=1)
    dec = _decoder(H, p, learned_gamma=net, learned_legs=1)
    dec.record_learned_gammas = True
    print("[gpu] learned inference", flush=True)
    dec.decode(torch.from_numpy(syn).cuda())
    torch.cuda.synchronize()
    hard = tr["hard"].cpu().numpy()[:, 0]
    reached = hard[:, 0] != 255
    ok = np.zeros_like(reached)
    ok[reached] = learned._syndrome_ok(H.tocsr().astype(np.int32), hard[:, None, :], syn)[reached, 0]
    u = dec.gamma_table(syn.shape[0]).cpu().numpy()[0]             # leg 0's draw
    feats = tr["features"].cpu().numpy()[:, 0][ok]                 # [n, V, 8]
    got = dec.learned_gammas.cpu().numpy()[:, 0]
    d = mo.shift_q8(net.params.cpu().numpy(), feats.reshape(-1, 8)).reshape(feats.shape[:2])
    exp = np.clip(u[None, :] + d, np.float32(-0.24), np.float32(0.66)).astype(np.float32)
    assert ok.sum() > 10 and np.abs(d).max() > 0.01
    np.testing.assert_allclose(got[ok], exp, atol=4e-3)            # one int8 step
    bad = reached & ~ok
    if bad.any():
        np.testing.assert_array_equal(got[bad], np.broadcast_to(u, (int(bad.sum()), u.size)))


@pytest.mark.gpu
def test_gpu_zero_network_reproduces_the_uniform_legs():
    """A network with every parameter 0 outputs z = 0: no shift, and the
    learned decode is the plain decode bit for bit."""
    H, p, err, syn = _banded(96, seed=14)
    S = torch.from_numpy(syn).cuda()
    base = _decoder(H, p).decode(S)
    dec = _decoder(H, p, learned_gamma=np.zeros(P, np.float32), learned_legs=3)
    print("[gpu] zero network", flush=True)
    out = dec.decode(S)
    torch.cuda.synchronize()
    for k in ("errors", "iterations", "found_count", "weights"):
        assert torch.equal(out[k], base[k]), k


@pytest.mark.gpu
def test_gpu_traces_record_the_state_at_each_leg():
    H, p, err, syn = _banded(48, seed=13)
    dec = _decoder(H, p)
    out = dec.collect_traces(torch.from_numpy(syn).cuda(), legs=2)
    torch.cuda.synchronize()
    f = out["features"].cpu().numpy()
    hard = out["hard"].cpu().numpy()
    V = H.shape[1]
    deg = np.diff(H.tocsc().indptr)
    reached = hard[:, 0, 0] != 255
    assert reached.any()
    np.testing.assert_allclose(f[reached, 0, :, 5], np.broadcast_to(deg / 8.0, (reached.sum(), V)))
    np.testing.assert_allclose(f[reached, 0, :, 7], 0.0)
    # hard decision = sign of the recorded posterior feature
    post = f[reached, 0, :, 1]
    np.testing.assert_array_equal(hard[reached, 0] == 1, post <= 0)
    # a shot that needed no relay leg records nothing
    done = out["found_count"].cpu().numpy() >= 4
    assert ((hard[:, 0, 0] == 255) <= done).all()


@pytest.mark.gpu
@pytest.mark.parametrize("pre_iter", [30, 0])
def test_gpu_leg0_features_match_the_oracle(pre_iter):
    """The feature pass at the start of relay leg 0 (gathered from the kernel's
    compressed check state) equals the oracle's restatement from its explicit
    per-edge messages, bit for bit, for every shot that reaches the leg: those
    whose pre-iter leg converged (feature 4 from one more check pass, feature 6
    zero), those whose did not (violated-check fractions), and with no pre-iter
    leg at all (every message still 0)."""
    from pelagos_amd import RelayBPDecoder

    H, p, err, syn = _banded(64, seed=12)
    dec = RelayBPDecoder(H, p, pre_iter=pre_iter, num_sets=12, set_max_iter=20, gamma0=0.65,
                         gamma_dist_interval=(-0.24, 0.66), stop_nconv=4, data_type="float32", algo="relay",
                         seed=3, device=[0])
    print(f"[gpu] leg-0 features against the oracle, pre_iter={pre_iter}", flush=True)
    tr = dec.collect_traces(torch.from_numpy(syn).cuda(), legs=1)
    torch.cuda.synchronize()
    feats = tr["features"].cpu().numpy()[:, 0]
    reached = tr["hard"].cpu().numpy()[:, 0, 0] != 255
    B, V = syn.shape[0], H.shape[1]
    lam = dec.log_prior_ratios_per_gpu[0].reshape(-1).cpu().numpy()
    rows = dec.gamma_table(B).cpu().numpy()
    kw = dict(pre_iter=pre_iter, num_sets=12, set_max_iter=20, stop_nconv=4, gamma0=0.65)
    ref = oracle.decode(dec.graph.chk_ptr, dec.graph.edge_var, V, syn, lam, rows, want_features0=True, **kw)
    pre = oracle.decode(dec.graph.chk_ptr, dec.graph.edge_var, V, syn, lam, rows, skip_relay_legs=True, **kw)
    np.testing.assert_array_equal(reached, pre["found_count"] < 4)
    conv = pre["found_count"] >= 1
    if pre_iter:
        assert (reached & conv).sum() > 10 and (reached & ~conv).sum() > 0
    else:
        assert reached.all() and not feats[..., 4].any()
    np.testing.assert_array_equal(feats[reached], ref["features0"][reached])


@pytest.mark.gpu
def test_gpu_leg_iterations_match_truncated_decoders():
    """Per-leg iteration counts (decodes stopped after r relay legs, on a fresh
    decoder that has drawn no gamma table yet) add up to the full decode's, and
    match decoders built with num_sets = r: every stopped decode draws a table
    of its own length."""
    H, p, err, syn = _banded(64, seed=17)
    s = torch.from_numpy(syn).cuda()
    dec = _decoder(H, p)
    its = learned.leg_iterations(dec, s, 3)
    full = dec.decode(s)["iterations"].cpu().numpy()
    first = _decoder(H, p).decode(s, skip_relay_legs=True)["iterations"].cpu().numpy()
    from pelagos_amd import RelayBPDecoder

    for r in (1, 3):
        dr = RelayBPDecoder(H, p, pre_iter=30, num_sets=r, set_max_iter=20, gamma0=0.65,
                            gamma_dist_interval=(-0.24, 0.66), stop_nconv=4, data_type="float32", algo="relay",
                            seed=3, device=[0])
        np.testing.assert_array_equal(first + its[:, :r].sum(1), dr.decode(s)["iterations"].cpu().numpy())
    assert (its >= 0).all() and (first + its.sum(1) <= full).all()
    assert dec.num_sets == 12


@pytest.mark.gpu
def test_gpu_trained_network_shortens_the_relay_legs():
    """What the shipped network gains on the 936 x 8785 circuit-shaped
    workload: the same shots take far fewer BP iterations than with the
    uniform legs (the relay legs that start from a valid solution converge in
    about two iterations instead of six), every claimed solution satisfies its
    syndrome, and the decoding agrees with the truth as often as the plain
    form's."""
    from pelagos_amd import RelayBPDecoder

    path = learned.default_weights_path()
    with np.load(path, allow_pickle=False) as f:
        legs = int(f["learned_legs"])
    H, p = codes.banded_circuit_code(seed=0)
    err = codes.sample_errors(p, 2048, seed=21)
    syn = codes.syndromes_of(H, err)
    S = torch.from_numpy(syn).cuda().contiguous()
    relay = dict(pre_iter=80, num_sets=100, set_max_iter=60, gamma0=0.65, gamma_dist_interval=(-0.24, 0.66),
                 stop_nconv=5, data_type="float32", algo="relay", seed=0, device=[0])
    plain = RelayBPDecoder(H, p, **relay).decode(S)
    dec = RelayBPDecoder(H, p, learned_gamma=learned.GammaMLP.load(path), learned_legs=legs, **relay)
    print("[gpu] trained network", flush=True)
    out = dec.decode(S)
    torch.cuda.synchronize()
    it_p = plain["iterations"].float().mean().item()
    it_l = out["iterations"].float().mean().item()
    assert it_l < 0.7 * it_p, (it_l, it_p)
    e = out["errors"].cpu().numpy()
    ok = out["found_count"].cpu().numpy() > 0
    assert (codes.syndromes_of(H, e)[ok] == syn[ok]).all()
    wrong_l = int((e != err).any(1).sum())
    wrong_p = int((plain["errors"].cpu().numpy() != err).any(1).sum())
    assert wrong_l <= wrong_p + 1, (wrong_l, wrong_p)


def test_last_damp_outside_the_unit_interval_is_rejected():
    H, p, err, syn = _banded(4, seed=1)
    for bad in (-0.1, 1.5):
        with pytest.raises(ValueError):
            _decoder(H, p, learned_last_damp=bad)
