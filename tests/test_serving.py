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
