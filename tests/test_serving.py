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


def test_result_arena_never_hands_out_rows_of_unfinished_requests():
    """The result arena's ring allocation, driven by random request sizes and
    completions: a range it hands out never overlaps a request still in
    flight, and it refuses (None) only when no such range exists."""
    import collections
    import types

    from pelagos_amd.serving import _Arena

    rng = np.random.default_rng(0)
    a = _Arena.__new__(_Arena)
    a.rows, a.head, a.live = 64, 0, collections.deque()
    inflight = []
    refused = 0
    for _ in range(5000):
        if inflight and rng.random() < 0.45:
            r = inflight.pop(int(rng.integers(len(inflight))))
            r.done = 1.0
            continue
        b = int(rng.integers(1, 20))
        r = types.SimpleNamespace(done=-1.0)
        o = a.take(b, r)
        busy = np.zeros(a.rows, bool)
        for q in inflight:
            busy[q.o:q.o + q.b] = True
        if o is None:
            refused += 1
            continue
        assert 0 <= o and o + b <= a.rows and not busy[o:o + b].any()
        if rng.random() < 0.1:          # a request refused after its rows were taken gives them back
            a.undo()
            continue
        r.o, r.b = o, b
        inflight.append(r)
    assert refused > 0


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
    rng = np.random.default_rng(1)
    pools = {}
    for k, (_, H, p) in dd.items():
        err = codes.sample_errors(p, 1024, seed=int(rng.integers(1 << 30)), scale=1.5)
        pools[k] = torch.from_numpy(codes.syndromes_of(H, err)).cuda()
    torch.cuda.synchronize()
    print(f"[gpu] serving {mode}", flush=True)
    with DecodeServer({k: v[0] for k, v in dd.items()}, max_batch=512) as server:
        assert server.roles()["bb72"]["group"] == 64 and server.roles()["banded4"]["group"] == 1024
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
        server.quiesce()
        assert reqs and all(r.done >= r.arrival for r in reqs)
        for r in reqs:
            dec = dd[r.code][0]
            alone = dec.decode(r.syndromes)
            for k, got in (("errors", r.errors), ("weights", r.weights), ("iterations", r.iterations),
                           ("found_count", r.found_count)):
                assert torch.equal(got, alone[k]), (r.rid, r.code, k)


@pytest.mark.gpu
def test_requests_join_a_running_launch_and_pass_a_slow_shot():
    """Shots that never converge (the full schedule, hundreds of ms on one
    workgroup) are submitted first; small requests submitted after them join
    the same launch and finish while the slow shots still run.  Every result
    equals the request decoded alone."""
    from pelagos_amd import RelayBPDecoder
    from pelagos_amd.serving import DecodeServer

    relay = dict(pre_iter=80, num_sets=100, set_max_iter=60, gamma0=0.65, gamma_dist_interval=(-0.24, 0.66),
                 stop_nconv=5, data_type="float32", algo="relay", seed=0, device=[0])
    Hs, ps = codes.banded_circuit_code(seed=0)                 # 936 x 8785
    Hb, pb = codes.bb_72_12_6(), np.full(72, 0.03)
    decs = {"slow": RelayBPDecoder(Hs, ps, **relay), "bb72": RelayBPDecoder(Hb, pb, **relay)}
    # heavy syndromes: about a third of the checks violated, far beyond what the code corrects
    slow_syn = torch.from_numpy((np.random.default_rng(4).random((4, Hs.shape[0])) < 0.35).astype(np.uint8)).cuda()
    pool = torch.from_numpy(codes.syndromes_of(Hb, codes.sample_errors(pb, 512, seed=5))).cuda()
    torch.cuda.synchronize()
    print("[gpu] serving: small requests past slow shots", flush=True)
    with DecodeServer(decs, max_batch=512, idle_ms=5.0) as server:
        slow = server.submit("slow", slow_syn)
        small = [server.submit("bb72", pool[32 * i:32 * i + 32]) for i in range(16)]
        server.drain_requests(small)
        passed = slow.done < 0
        launches = server.launches
        server.drain()
        server.quiesce()
        full = 80 + 100 * 60
        print(f"[gpu] slow shots: iterations {slow.iterations.tolist()}, launches {launches}", flush=True)
        assert int(slow.iterations.max()) == full   # the premise: a shot that runs the whole schedule
        assert passed, "the slow shots finished before the small requests"
        assert launches == 1, launches
        for r in [slow] + small:
            alone = decs[r.code].decode(r.syndromes)
            for k, got in (("errors", r.errors), ("weights", r.weights), ("iterations", r.iterations),
                           ("found_count", r.found_count)):
                assert torch.equal(got, alone[k]), (r.rid, r.code, k)


@pytest.mark.gpu
def test_a_request_completes_after_its_ring_entries_are_reused():
    """A request whose first unit finishes at once and whose second runs the
    whole schedule, while more than the ring's capacity of later units pass
    through (the first unit's entry is reused meanwhile): the request still
    completes, with the results of decoding it alone."""
    from pelagos_amd import RelayBPDecoder
    from pelagos_amd.serving import DecodeServer, ServerBusy

    relay = dict(pre_iter=80, num_sets=100, set_max_iter=60, gamma0=0.65, gamma_dist_interval=(-0.24, 0.66),
                 stop_nconv=5, data_type="float32", algo="relay", seed=0, device=[0])
    Hs, ps = codes.banded_circuit_code(seed=0)
    Hb, pb = codes.bb_72_12_6(), np.full(72, 0.03)
    decs = {"slow": RelayBPDecoder(Hs, ps, **relay), "bb72": RelayBPDecoder(Hb, pb, **relay)}
    heavy = (np.random.default_rng(4).random(Hs.shape[0]) < 0.35).astype(np.uint8)
    pair = torch.from_numpy(np.stack([np.zeros_like(heavy), heavy])).cuda()
    pool = torch.from_numpy(codes.syndromes_of(Hb, codes.sample_errors(pb, 256, seed=8))).cuda()
    torch.cuda.synchronize()
    print("[gpu] serving: ring entries reused under a pending request", flush=True)
    with DecodeServer(decs, max_batch=64, max_queued_shots=64, queue_units=64, idle_ms=5.0) as server:
        r = server.submit("slow", pair)
        small, i = [], 0
        while i < 100:
            try:
                small.append(server.submit("bb72", pool[i:i + 1]))
                i += 1
            except ServerBusy:
                server.poll()
        server.drain(timeout_s=60.0)
        server.quiesce()
        assert small[-1].first_unit >= 64 + 2
        for q in [r] + small:
            alone = decs[q.code].decode(q.syndromes)
            for k, got in (("errors", q.errors), ("weights", q.weights), ("iterations", q.iterations),
                           ("found_count", q.found_count)):
                assert torch.equal(got, alone[k]), (q.rid, q.code, k)


@pytest.mark.gpu
def test_backpressure_and_a_full_queue():
    """Requests beyond max_queued_shots are refused with ServerBusy and admitted
    again once earlier ones finish; the C queue refuses (PEL_EBUSY) a unit whose
    entry still holds an unfinished one."""
    import ctypes

    from pelagos_amd import _native
    from pelagos_amd.serving import DecodeServer, ServerBusy

    dd = _decoders()
    H, p = dd["bb72"][1], dd["bb72"][2]
    pool = torch.from_numpy(codes.syndromes_of(H, codes.sample_errors(p, 256, seed=6))).cuda()
    torch.cuda.synchronize()
    print("[gpu] serving backpressure", flush=True)
    with DecodeServer({"bb72": dd["bb72"][0]}, max_batch=64, max_queued_shots=128) as server:
        a = server.submit("bb72", pool[:64])
        b = server.submit("bb72", pool[64:128])
        with pytest.raises(ServerBusy):
            server.submit("bb72", pool[128:129])
        server.drain()
        c = server.submit("bb72", pool[128:192])
        server.drain()
        server.quiesce()
        alone = dd["bb72"][0].decode(pool[:192])
        got = torch.cat([r.errors for r in (a, b, c)])
        assert torch.equal(got, alone["errors"])

    L = _native.lib()
    q = ctypes.c_void_p()
    _native.check(L.pel_serve_queue_create(2, ctypes.byref(q)), "create")
    try:
        buf = torch.zeros(4096, dtype=torch.uint8, device="cuda")
        u = _native.PelServeUnit(job=0, count=1, syn_row_stride=36, syn_col_stride=1, syndromes=buf.data_ptr(),
                                 errors=buf.data_ptr(),
                                 weights=buf.data_ptr(), iterations=buf.data_ptr(), found_count=buf.data_ptr())
        units = (_native.PelServeUnit * 2)(u, u)
        first = ctypes.c_int64()
        assert L.pel_serve_publish(q, units, 2, ctypes.byref(first)) == 0 and first.value == 0
        assert L.pel_serve_publish(q, units, 1, ctypes.byref(first)) == _native.PEL_EBUSY
        bad = (_native.PelServeUnit * 1)(_native.PelServeUnit(job=0, count=65, syndromes=buf.data_ptr()))
        assert L.pel_serve_publish(q, bad, 1, None) == _native.PEL_EINVAL
        pub = ctypes.c_int64()
        assert L.pel_serve_queue_info(q, None, None, ctypes.byref(pub)) == 0 and pub.value == 2
    finally:
        L.pel_serve_queue_destroy(q)


@pytest.mark.gpu
def test_hbm_budget_sizes_the_server_and_bounds_admission():
    """With hbm_budget_bytes the server holds at most that much device memory:
    the result arenas take the rows the budget leaves after the fixed part, the
    admission bound is a quarter of them, requests beyond it are refused, and a
    budget that cannot hold one full request is refused."""
    from pelagos_amd.serving import DecodeServer, ServerBusy

    dd = _decoders()
    dec, H, p = dd["bb72"]
    pool = torch.from_numpy(codes.syndromes_of(H, codes.sample_errors(p, 512, seed=8))).cuda()
    torch.cuda.synchronize()
    print("[gpu] serving hbm budget", flush=True)
    with DecodeServer({"bb72": dec}, max_batch=64) as probe:
        fixed = probe.device_bytes() - probe.result_shots * (dec.V + 12)
    row = dec.V + 12
    budget = fixed + 1024 * row + row // 2
    with DecodeServer({"bb72": dec}, max_batch=64, hbm_budget_bytes=budget) as server:
        assert server.result_shots == 1024 and server.max_queued_shots == 256
        assert server.device_bytes() <= budget
        reqs = [server.submit("bb72", pool[64 * i:64 * (i + 1)]) for i in range(4)]
        with pytest.raises(ServerBusy):
            server.submit("bb72", pool[256:257])
        server.drain()
        reqs.append(server.submit("bb72", pool[256:320]))
        server.drain()
        server.quiesce()
        alone = dec.decode(pool[:320])
        assert torch.equal(torch.cat([r.errors for r in reqs]), alone["errors"])
    with pytest.raises(ValueError):
        DecodeServer({"bb72": dec}, max_batch=64, hbm_budget_bytes=fixed + 63 * row)
