"""pel_decode_noalloc (the allocation-free decode) against pel_decode, through
the C ABI: two calls in a row on each workload, the second with fewer shots,
bit for bit on every output; and a decode too large for the library's
per-GPU workspace is refused before anything is launched."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from pelagos_amd import _native, codes, learned

pytestmark = pytest.mark.gpu

RELAY = dict(pre_iter=80, num_sets=100, set_max_iter=60, gamma0=0.65, gamma_dist_interval=(-0.24, 0.66),
             stop_nconv=5, data_type="float32", algo="relay", seed=0, device=[0])
KEYS = ("errors", "weights", "iterations", "found_count")


def _buffers(dec, B):
    dev = dec.device
    return dict(errors=torch.full((B, dec.V), 7, dtype=torch.uint8, device=dev),
                weights=torch.full((B,), -1.0, dtype=torch.float32, device=dev),
                iterations=torch.full((B,), -1, dtype=torch.int32, device=dev),
                found_count=torch.full((B,), -1, dtype=torch.int32, device=dev))


def _noalloc(dec, syn):
    B = int(syn.shape[0])
    dec.prepare(B)
    out = _buffers(dec, B)
    s = torch.cuda.current_stream(dec.device)
    dec.decode_async(syn, out["errors"], out["weights"], out["iterations"], out["found_count"], None, s)
    torch.cuda.synchronize()
    return out


def _workloads():
    w = codes.three_code_workload(0.003)
    H1, p1 = codes.banded_circuit_code(seed=0)
    path = learned.default_weights_path()
    with np.load(path, allow_pickle=False) as f:
        legs = int(f["learned_legs"])
    net = dict(learned_gamma=learned.GammaMLP.load(path), learned_legs=legs)
    return [("banded936", H1, p1, {}, (256, 64)),
            ("banded936-learned", H1, p1, net, (128, 32)),
            ("banded1872", *w["banded1872"], {}, (64, 16)),
            ("bb72", *w["bb72"], {}, (1024, 100))]


@pytest.mark.parametrize("case", range(4), ids=["banded936", "banded936-learned", "banded1872", "bb72"])
def test_noalloc_decode_equals_pel_decode_twice(case):
    from pelagos_amd import RelayBPDecoder

    name, H, p, kw, sizes = _workloads()[case]
    dec = RelayBPDecoder(H, p, **RELAY, **kw)
    err = codes.sample_errors(p, sizes[0], seed=41)
    syn_all = torch.from_numpy(codes.syndromes_of(H, err)).cuda()
    for B in sizes:                                   # the second call has fewer shots
        syn = syn_all[sizes[0] - B:].contiguous()     # other shots than the first call's prefix
        plan = dec.plan(B, kw.get("learned_gamma") is not None)
        print(f"[gpu] noalloc {name} B={B} placement={plan.placement} requeue={plan.requeue_below}", flush=True)
        got = _noalloc(dec, syn)
        ref = dec.decode(syn)
        torch.cuda.synchronize()
        for k in KEYS:
            assert torch.equal(got[k], ref[k]), (name, B, k)


def test_noalloc_refuses_a_decode_beyond_its_workspace():
    """4096 shots of the 936 x 8785 code with re-queue on park [4096, 8785]
    posteriors (144 MB): refused with a ValueError naming the limit, and the
    next call decodes normally."""
    from pelagos_amd import RelayBPDecoder

    H, p = codes.banded_circuit_code(seed=0)
    dec = RelayBPDecoder(H, p, **RELAY)
    B = 4096
    dec.prepare(B)
    assert dec.plan(B).workspace_bytes > _native.PEL_NOALLOC_MAX_BYTES
    syn = torch.zeros(B, dec.C, dtype=torch.uint8, device="cuda")
    out = _buffers(dec, B)
    with pytest.raises(ValueError, match="bytes of workspace"):
        dec.decode_async(syn, out["errors"], out["weights"], out["iterations"], out["found_count"], None,
                         torch.cuda.current_stream())
    assert (out["iterations"] == -1).all()            # nothing was launched
    small = torch.from_numpy(codes.syndromes_of(H, codes.sample_errors(p, 32, seed=5))).cuda().contiguous()
    got = _noalloc(dec, small)
    ref = dec.decode(small)
    for k in KEYS:
        assert torch.equal(got[k], ref[k]), k
