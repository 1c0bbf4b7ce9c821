"""Several codes and batches in one persistent launch (pel_decode_jobs), also
captured in a graph, and the first-sight launch-shape tuning of decode().

Every job's results are checked against an ordinary decode() of the same
syndromes, bit for bit; the graph is replayed several times with fresh
syndromes and eager decodes in between."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from pelagos_amd import codes

pytestmark = pytest.mark.gpu

RELAY = dict(pre_iter=40, num_sets=20, set_max_iter=30, gamma0=0.65, gamma_dist_interval=(-0.24, 0.66),
             stop_nconv=3, data_type="float32", algo="relay", seed=9, device=[0])


def _codes():
    Hs, ps = codes.banded_circuit_code(n_blocks=4, V=2700, seed=2)
    Hm, pm = codes.banded_circuit_code(n_blocks=6, V=4000, seed=5)
    return {"bb72": (codes.bb_72_12_6(), np.full(72, 0.03)), "banded4": (Hs, ps), "banded6": (Hm, pm)}


def test_graph_replays_match_eager_decodes():
    from pelagos_amd import RelayBPDecoder
    from pelagos_amd.batching import MultiCodeBatch

    cs = _codes()
    decs = {k: RelayBPDecoder(H, p, **RELAY) for k, (H, p) in cs.items()}
    shots = {"bb72": 512, "banded4": 128, "banded6": 64}
    mb = MultiCodeBatch(decs, shots)
    print("[gpu] capture", flush=True)
    mb.capture()
    rng = np.random.default_rng(0)
    for rep in range(3):
        syn = {}
        for k, (H, p) in cs.items():
            err = codes.sample_errors(p, shots[k], seed=int(rng.integers(1 << 30)), scale=1.5)
            syn[k] = torch.from_numpy(codes.syndromes_of(H, err)).cuda()
            mb.syndromes[k].copy_(syn[k])
        print(f"[gpu] replay {rep}", flush=True)
        mb.replay()
        torch.cuda.synchronize()
        for k, dec in decs.items():
            ref = dec.decode(syn[k])          # eager work between replays
            torch.cuda.synchronize()
            for key in ("errors", "weights", "iterations", "found_count"):
                assert torch.equal(mb.out[k][key], ref[key]), (rep, k, key)


def test_decode_tunes_its_launch_shape_on_first_sight(tmp_path, monkeypatch):
    """autotune=True: the first decode of a batch size times the candidate
    launch shapes (block size, re-queue threshold, state placement), keeps the
    fastest and writes it to the cache; the results
    equal an untuned decoder's, and a second decoder of the same problem takes
    the shape from the cache without timing anything."""
    from pelagos_amd import RelayBPDecoder
    from pelagos_amd.decoder import tune_cache_path

    monkeypatch.setenv("PELAGOS_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    H, p = codes.banded_circuit_code(n_blocks=4, V=2700, seed=2)
    syn = torch.from_numpy(codes.syndromes_of(H, codes.sample_errors(p, 256, seed=1, scale=1.5))).cuda()
    base = RelayBPDecoder(H, p, autotune=False, **RELAY).decode(syn)
    dec = RelayBPDecoder(H, p, autotune=True, **RELAY)
    print("[gpu] autotune", flush=True)
    out = dec.decode(syn)
    assert dec.tune_runs == 1 and tune_cache_path().exists()
    blk, rq, pl = dec._tuned[(256, False)]
    assert blk in (256, 512, 1024) and rq in (0, 16, 32) and pl in (15, 7)
    assert dec.plan(256).block == blk and dec.plan(256).requeue_below == rq and dec.plan(256).placement == pl
    for key in ("errors", "weights", "iterations", "found_count"):
        assert torch.equal(out[key], base[key]), key
    again = RelayBPDecoder(H, p, autotune=True, **RELAY)
    out2 = again.decode(syn)
    assert again.tune_runs == 0 and again._tuned[(256, False)] == (blk, rq, pl)
    for key in ("errors", "weights", "iterations", "found_count"):
        assert torch.equal(out2[key], base[key]), key


def test_decode_async_matches_decode_and_checks_its_buffers():
    """decode_async on side streams with caller-owned buffers gives decode()'s
    results; buffers the kernel cannot take as raw pointers are refused."""
    from pelagos_amd import RelayBPDecoder

    H, p = _codes()["banded4"]
    dec = RelayBPDecoder(H, p, **RELAY)
    err = codes.sample_errors(p, 96, seed=4, scale=1.5)
    raw = torch.from_numpy(codes.syndromes_of(H, err)).cuda()   # transposed layout: not contiguous
    syn = raw.contiguous()
    ref = dec.decode(syn)
    B, V = syn.shape[0], dec.V
    ws = torch.empty(dec.prepare(B), dtype=torch.uint8, device="cuda")
    out = dict(errors=torch.empty(B, V, dtype=torch.uint8, device="cuda"),
               weights=torch.empty(B, dtype=torch.float32, device="cuda"),
               iterations=torch.empty(B, dtype=torch.int32, device="cuda"),
               found_count=torch.empty(B, dtype=torch.int32, device="cuda"))
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    dec.decode_async(syn, out["errors"], out["weights"], out["iterations"], out["found_count"], ws, s)
    torch.cuda.synchronize()
    for k, v in out.items():
        assert torch.equal(v, ref[k]), k
    if not raw.is_contiguous():
        with pytest.raises(ValueError, match="contiguous"):
            dec.decode_async(raw, out["errors"], out["weights"], out["iterations"], out["found_count"], ws, s)
    with pytest.raises(ValueError, match="iterations"):
        dec.decode_async(syn, out["errors"], out["weights"], out["iterations"].long(), out["found_count"], ws, s)


def _check_jobs(jl, syn, decs):
    for name, dec in decs.items():
        if jl.shots[name] == 0:
            continue
        ref = dec.decode(syn[name])
        torch.cuda.synchronize()
        for key in ("errors", "weights", "iterations", "found_count"):
            assert torch.equal(jl.out[name][key], ref[key]), (name, key)


def test_one_launch_over_mixed_jobs_equals_separate_decodes():
    """One pel_decode_jobs launch over the [[72,12,6]] code (a shot per wave),
    two banded codes (a shot per workgroup), a learned-strength job and an
    empty job: every job's results equal its own decode(), on two launches
    with fresh syndromes."""
    from pelagos_amd import RelayBPDecoder, learned
    from pelagos_amd.batching import JobLaunch

    cs = _codes()
    H4, p4 = cs["banded4"]
    decs = {k: RelayBPDecoder(H, p, **RELAY) for k, (H, p) in cs.items()}
    decs["learned4"] = RelayBPDecoder(H4, p4, learned_gamma=learned.GammaMLP(seed=2), learned_legs=2, **RELAY)
    decs["empty"] = RelayBPDecoder(H4, p4, **RELAY)
    shots = {"bb72": 300, "banded4": 50, "empty": 0, "learned4": 40, "banded6": 20}
    jl = JobLaunch([(k, decs[k], shots[k]) for k in ("bb72", "banded4", "empty", "learned4", "banded6")])
    roles = jl.roles()
    assert roles["bb72"]["group"] == 64 and roles["banded4"]["group"] == 1024 and roles["learned4"]["learned"]
    rng = np.random.default_rng(3)
    for rep in range(2):
        syn = {}
        for k, dec in decs.items():
            H, p = (H4, p4) if k in ("learned4", "empty") else cs[k]
            err = codes.sample_errors(p, shots[k], seed=int(rng.integers(1 << 30)), scale=1.5)
            syn[k] = torch.from_numpy(codes.syndromes_of(H, err)).cuda().contiguous()
            jl.syndromes[k].copy_(syn[k])
        print(f"[gpu] job launch rep {rep}", flush=True)
        jl.launch()
        torch.cuda.synchronize()
        _check_jobs(jl, syn, decs)


def test_one_launch_over_batches_of_one_code():
    """Five batches of the 936 x 8785 code in one launch (the later batches fill
    the earlier ones' tails): each equals its own decode()."""
    from pelagos_amd import RelayBPDecoder
    from pelagos_amd.batching import JobLaunch

    H, p = codes.banded_circuit_code(seed=0)
    dec = RelayBPDecoder(H, p, **dict(RELAY, pre_iter=80, num_sets=100, set_max_iter=60, stop_nconv=5, seed=0))
    jl = JobLaunch([(f"b{i}", dec, 96) for i in range(5)])
    syn = {}
    for i in range(5):
        err = codes.sample_errors(p, 96, seed=60 + i)
        syn[f"b{i}"] = torch.from_numpy(codes.syndromes_of(H, err)).cuda().contiguous()
        jl.syndromes[f"b{i}"].copy_(syn[f"b{i}"])
    print("[gpu] five batches in one launch", flush=True)
    jl.launch()
    torch.cuda.synchronize()
    _check_jobs(jl, syn, {f"b{i}": dec for i in range(5)})


@pytest.mark.gpu
def test_waves_go_on_across_batches_of_a_small_code():
    """Four batches of the [[72,12,6]] code decoded a shot per wave, an empty
    batch among them and a 936 x 8785 batch after them, in one launch: the
    waves move from batch to batch on their own, and every batch equals its own
    decode()."""
    from pelagos_amd import RelayBPDecoder
    from pelagos_amd.batching import JobLaunch

    Hb = codes.bb_72_12_6()
    pb = np.full(72, 0.03)
    small = RelayBPDecoder(Hb, pb, **RELAY)
    H, p = codes.banded_circuit_code(seed=0)
    big = RelayBPDecoder(H, p, **RELAY)
    jobs = [("s0", small, 700), ("s1", small, 300), ("e", small, 0), ("s2", small, 513), ("s3", small, 64),
            ("big", big, 40)]
    jl = JobLaunch(jobs)
    assert jl.roles()["s0"]["group"] == 64 and jl.roles()["big"]["group"] == 1024
    syn = {}
    for i, (name, dec, n) in enumerate(jobs):
        HH, pp = (Hb, pb) if dec is small else (H, p)
        err = codes.sample_errors(pp, n, seed=80 + i, scale=1.5)
        syn[name] = torch.from_numpy(codes.syndromes_of(HH, err)).cuda().contiguous()
        jl.syndromes[name].copy_(syn[name])
    print("[gpu] waves across small-code batches", flush=True)
    jl.launch()
    torch.cuda.synchronize()
    _check_jobs(jl, syn, {name: dec for name, dec, _ in jobs})
