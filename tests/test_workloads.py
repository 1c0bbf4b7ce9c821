"""The three codes of the batched and serving workloads on the GPU, against the
oracle: the 1872 x 17570 banded code at its own plan (its state does not fit
LDS, so the plan spills it), the learned form on the 936 x 8785 code with the
shipped network, the three codes in one captured graph, and the server over
all three codes with the learned form on the 936 x 8785 code.

Bar: bit-exact errors, iterations, found counts and weights (and posteriors
where the oracle is asked for them)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from pelagos_amd import codes, learned

pytestmark = pytest.mark.gpu

RELAY = dict(pre_iter=80, num_sets=100, set_max_iter=60, gamma0=0.65, gamma_dist_interval=(-0.24, 0.66),
             stop_nconv=5)
KEYS = ("errors", "iterations", "found_count", "weights")


def _decoder(H, p, **kw):
    from pelagos_amd import RelayBPDecoder

    return RelayBPDecoder(H, p, data_type="float32", algo="relay", seed=0, device=[0], **RELAY, **kw)


def _oracle(dec, syn, table=None, want_posterior=False):
    """The oracle on dec's graph, priors and parameters; ``table`` replaces the
    decoder's gamma rows ([R, V], or [R, B, V] per shot)."""
    B = syn.shape[0]
    tab = dec.gamma_table(B).cpu().numpy() if table is None else table
    lam = dec.log_prior_ratios_per_gpu[0].reshape(-1).cpu().numpy()
    return oracle.decode(dec.graph.chk_ptr, dec.graph.edge_var, dec.V, syn, lam, tab,
                         pre_iter=RELAY["pre_iter"], num_sets=RELAY["num_sets"],
                         set_max_iter=RELAY["set_max_iter"], stop_nconv=RELAY["stop_nconv"],
                         gamma0=RELAY["gamma0"], gamma_per_row=tab.ndim == 3, want_posterior=want_posterior)


def _same(out, ref, keys=KEYS, tag=""):
    for k in keys:
        got = out[k].cpu().numpy() if torch.is_tensor(out[k]) else out[k]
        if not np.array_equal(got, ref[k]):
            bad = np.nonzero((got != ref[k]).reshape(got.shape[0], -1).any(1))[0]
            raise AssertionError(f"{tag} {k}: {bad.size} shots differ, first {bad[:8].tolist()}")


def _learned_kw():
    path = learned.default_weights_path()
    with np.load(path, allow_pickle=False) as f:
        legs = int(f["learned_legs"])
        damp = float(f["learned_last_damp"]) if "learned_last_damp" in f.files else 0.0
    return dict(learned_gamma=learned.GammaMLP.load(path), learned_legs=legs, learned_last_damp=damp)


def _learned_table(dec, B):
    """Per-shot gamma rows [R, B, V]: the strengths each learned leg used (as
    recorded by the decode), then the uniform rows."""
    gl = dec.learned_gammas.cpu().numpy()                      # [B, L, V]
    rows = dec.gamma_table(B).cpu().numpy()                    # [R, V]
    table = np.broadcast_to(rows[:, None, :], (rows.shape[0], B, dec.V)).copy()
    table[:gl.shape[1]] = np.transpose(gl, (1, 0, 2))
    return table


def test_banded1872_at_its_own_plan_matches_oracle():
    """64 shots of the 1872 x 17570 code at mean prior 0.003 with the full
    schedule; its per-shot state (over 140 KB of variable state alone) is
    planned with the check state and sign words spilled to the workspace."""
    H, p = codes.three_code_workload(0.003)["banded1872"]
    err = codes.sample_errors(p, 64, seed=31)
    syn = codes.syndromes_of(H, err)
    dec = _decoder(H, p)
    plan = dec.plan(64)
    assert (plan.placement & 7) != 7, plan.placement
    print(f"[gpu] banded1872 B=64 placement={plan.placement} block={plan.block}", flush=True)
    out = dec.decode(torch.from_numpy(syn).cuda(), return_last_posterior=True)
    torch.cuda.synchronize()
    ref = _oracle(dec, syn, want_posterior=True)
    _same(out, ref, KEYS + ("last_posterior",), "banded1872")
    solved = np.isfinite(ref["weights"])
    assert solved.mean() > 0.9
    assert (codes.syndromes_of(H, out["errors"].cpu().numpy()[solved]) == syn[solved]).all()


def test_learned_form_on_banded936_matches_oracle_given_its_strengths():
    """The shipped network on the 936 x 8785 code (64 shots, its own priors):
    the decode equals the oracle fed the strengths the network gave."""
    H, p = codes.banded_circuit_code(seed=0)
    err = codes.sample_errors(p, 64, seed=21)
    syn = codes.syndromes_of(H, err)
    dec = _decoder(H, p, **_learned_kw())
    dec.record_learned_gammas = True
    print("[gpu] learned banded936 B=64", flush=True)
    out = dec.decode(torch.from_numpy(syn).cuda())
    torch.cuda.synchronize()
    table = _learned_table(dec, 64)
    assert (table[:dec.learned_legs] != dec.gamma_table(64).cpu().numpy()[:dec.learned_legs, None, :]).any()
    _same(out, _oracle(dec, syn, table), tag="learned banded936")


def test_three_codes_in_one_graph_match_oracle_checked_decodes():
    """The three codes at mean prior 0.003 (512 / 128 / 32 shots) captured as
    one graph, replayed twice with fresh syndromes: every replay equals eager
    decodes of the same syndromes, which equal the oracle."""
    from pelagos_amd.batching import MultiCodeBatch

    cs = codes.three_code_workload(0.003)
    shots = {"bb72": 512, "banded936": 128, "banded1872": 32}
    decs = {k: _decoder(H, p) for k, (H, p) in cs.items()}
    mb = MultiCodeBatch(decs, shots)
    print("[gpu] three-code graph capture", flush=True)
    mb.capture()
    for rep in range(2):
        syn = {}
        for i, (k, (H, p)) in enumerate(cs.items()):
            syn[k] = codes.syndromes_of(H, codes.sample_errors(p, shots[k], seed=100 * rep + i))
            mb.syndromes[k].copy_(torch.from_numpy(syn[k]))
        print(f"[gpu] three-code replay {rep}", flush=True)
        mb.replay()
        torch.cuda.synchronize()
        for k, dec in decs.items():
            eager = dec.decode(torch.from_numpy(syn[k]).cuda())
            torch.cuda.synchronize()
            for key in KEYS:
                assert torch.equal(mb.out[k][key], eager[key]), (rep, k, key)
            if rep == 0:
                _same(eager, _oracle(dec, syn[k]), tag=k)


def test_server_over_three_codes_with_learned_form():
    """The server over the three codes (learned form on the 936 x 8785 code)
    with a burst of requests of 1-96 shots: each result equals decoding the
    request alone, and a request of the 936 x 8785 code decoded alone equals
    the oracle fed the strengths its network gave."""
    from pelagos_amd.serving import DecodeServer

    cs = codes.three_code_workload(0.003)
    decs = {k: _decoder(H, p, **(_learned_kw() if k == "banded936" else {})) for k, (H, p) in cs.items()}
    pools = {k: codes.syndromes_of(H, codes.sample_errors(p, 512, seed=7 + i)) for i, (k, (H, p)) in
             enumerate(cs.items())}
    subs = []
    rng = np.random.default_rng(3)
    for i in range(18):
        k = codes.THREE_CODES[i % 3]
        b = int(rng.integers(1, 97))
        o = int(rng.integers(0, 512 - b))
        subs.append((k, torch.from_numpy(pools[k][o:o + b]).cuda()))
    torch.cuda.synchronize()
    print("[gpu] three-code server burst", flush=True)
    with DecodeServer(decs, max_batch=512, max_queued_shots=1 << 16) as server:
        reqs = [server.submit(k, syn) for k, syn in subs]
        server.drain()
        server.quiesce()
    for r in reqs:
        alone = decs[r.code].decode(r.syndromes)
        for k, got in (("errors", r.errors), ("weights", r.weights), ("iterations", r.iterations),
                       ("found_count", r.found_count)):
            assert torch.equal(got, alone[k]), (r.rid, r.code, k)
    r = next(q for q in reqs if q.code == "banded936" and q.syndromes.shape[0] >= 8)
    dec = decs["banded936"]
    dec.record_learned_gammas = True
    out = dec.decode(r.syndromes)
    torch.cuda.synchronize()
    dec.record_learned_gammas = False
    B = r.syndromes.shape[0]
    _same(out, _oracle(dec, r.syndromes.cpu().numpy(), _learned_table(dec, B)), tag="served banded936")
