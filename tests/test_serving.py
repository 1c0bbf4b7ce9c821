"""Continuous-batching server: every request's result equals decoding it
alone (bit for bit), whether requests arrive in a burst or over time, and
the trace generator has the stated mean rate."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from pelagos_amd import codes
from pelagos_amd.serving import poisson_trace


def test_poisson_trace_rate_and_sizes():
    tr = poisson_trace(50_000, 4.0, ["a", "b", "c"], seed=0, max_shots=256)
    shots = sum(b for _, _, b in tr)
    assert abs(shots / 4.0 - 50_000) < 0.1 * 50_000
    assert all(1 <= b <= 256 for _, _, b in tr)
    assert all(t1 <= t2 for (t1, _, _), (t2, _, _) in zip(tr, tr[1:]))
    assert {c for _, c, _ in tr} == {"a", "b", "c"}


class _ScriptedServer:
    """Stands in for DecodeServer on the CPU: a queue of at most ``cap`` shots
    that completes one request per poll."""

    def __init__(self, cap):
        import collections
        import time

        self.cap, self.q, self.finished, self.queued_shots = cap, collections.deque(), [], 0
        self.t0 = time.perf_counter()

    def now(self):
        import time

        return time.perf_counter() - self.t0

    def submit(self, code, syn, arrival=None):
        from types import SimpleNamespace

        from pelagos_amd.serving import ServerBusy

        b = int(syn.shape[0])
        if self.queued_shots + b > self.cap:
            raise ServerBusy("full")
        r = SimpleNamespace(code=code, syndromes=syn, arrival=arrival, done=-1.0)
        self.q.append(r)
        self.queued_shots += b
        return r

    def poll(self):
        if self.q:
            r = self.q.popleft()
            self.queued_shots -= int(r.syndromes.shape[0])
            r.done = self.now()
            self.finished.append(r)

    def pending(self):
        return len(self.q)


@pytest.mark.parametrize("policy", ["reject", "defer"])
def test_admission_policy_accounts_for_every_request(policy):
    """A burst larger than the queue: under "reject" the refused requests are
    counted and dropped, under "defer" every request is served eventually, in
    arrival order, and none is refused."""
    from pelagos_amd.serving import run_trace

    trace = [(0.0, "a", 10)] * 20
    pools = {"a": torch.zeros(64, 3, dtype=torch.uint8)}
    srv = _ScriptedServer(cap=35)
    st = run_trace(srv, trace, pools, policy=policy)
    if policy == "reject":
        assert st["requests"] + st["rejected_requests"] == 20 and st["rejected_requests"] > 0
        assert st["rejected_shots"] == 10 * st["rejected_requests"] and st["deferred_requests"] == 0
    else:
        assert st["requests"] == 20 and st["rejected_requests"] == 0 and st["deferred_requests"] > 0
        assert [r.arrival for r in srv.finished] == sorted(r.arrival for r in srv.finished)


def _decoders():
    from pelagos_amd import RelayBPDecoder

    relay = dict(pre_iter=40, num_sets=20, set_max_iter=30, gamma0=0.65, gamma_dist_interval=(-0.24, 0.66),
                 stop_nconv=3, data_type="float32", algo="relay", seed=5, device=[0])
    Hb = codes.bb_72_12_6()
    Hs, ps = codes.banded_circuit_code(n_blocks=4, V=2700, seed=2)
    return {
        "bb72": (RelayBPDecoder(Hb, np.full(72, 0.03), **relay), Hb, np.full(72, 0.03)),
        "banded4": (RelayBPDecoder(Hs, ps, **relay), Hs, ps),
    }


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["burst", "trace"])
def test_requests_decode_the_same_alone_and_in_the_mix(mode):
    from pelagos_amd.serving import DecodeServer, run_trace

    dd = _decoders()
    server = DecodeServer({k: v[0] for k, v in dd.items()}, max_batch=512, slots=3)
    rng = np.random.default_rng(1)
    pools = {}
    for k, (_, H, p) in dd.items():
        err = codes.sample_errors(p, 1024, seed=int(rng.integers(1 << 30)), scale=1.5)
        pools[k] = torch.from_numpy(codes.syndromes_of(H, err)).cuda()
    print(f"[gpu] serving {mode}", flush=True)
    if mode == "burst":
        reqs = []
        for _ in range(24):
            k = ["bb72", "banded4"][int(rng.integers(2))]
            b = int(rng.integers(1, 97))
            o = int(rng.integers(0, 1024 - b))
            reqs.append(server.submit(k, pools[k][o:o + b]))
        server.drain()
    else:
        trace = poisson_trace(20_000, 0.3, ["bb72", "banded4"], seed=2, max_shots=64)
        stats = run_trace(server, trace, pools)
        assert stats["requests"] == len(trace) and stats["p99_ms"] >= stats["p50_ms"] > 0
        reqs = server.finished
    torch.cuda.synchronize()
    assert reqs and all(r.done >= r.arrival for r in reqs)
    for r in reqs:
        dec = dd[r.code][0]
        alone = dec.decode(r.syndromes)
        for k, got in (("errors", r.errors), ("weights", r.weights), ("iterations", r.iterations),
                       ("found_count", r.found_count)):
            assert torch.equal(got, alone[k]), (r.rid, r.code, k)
