"""Observable harness: the reference runner's result classes on the
[[72,12,6]] BB code, whose 12 logical X operators are the observables."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from pelagos_amd import codes
from pelagos_amd.observables import ObservableDecoderRunner, wilson


def test_wilson_interval_contains_the_rate_and_shrinks():
    p, lo, hi = wilson(30, 1000)
    assert lo < p < hi and abs(p - 0.03) < 1e-12
    _, lo2, hi2 = wilson(300, 10000)
    assert hi2 - lo2 < hi - lo
    assert wilson(0, 100)[1] == 0.0


class _FakeDecoder:
    """Decodes every syndrome to the all-zero pattern except shot 0 (true error)."""

    def __init__(self, H, truth):
        self.H_csr, self.truth = H, truth
        self.C, self.V = H.shape
        self.device = torch.device("cpu")
        self.bitpack_output = False
        self.wj = torch.ones(self.V)

    def decode(self, syn):
        B = syn.shape[0]
        e = torch.zeros(B, self.V, dtype=torch.uint8)
        e[0] = torch.from_numpy(self.truth[0])
        found = torch.tensor([5] + [0] * (B - 1), dtype=torch.int32)
        return dict(errors=e, iterations=torch.full((B,), 7, dtype=torch.int32), found_count=found,
                    unconverged_indices=torch.nonzero(found < 5).flatten().to(torch.int32))


def test_runner_classifies_shots_like_the_reference():
    H = codes.bb_72_12_6()
    L = codes.logical_operators(H.toarray(), codes.bb_72_12_6_hz())
    truth = codes.sample_errors(np.full(72, 0.05), 8, seed=3)
    truth[1] = 0                                   # shot 1: nothing to find, zero decoding is right
    run = ObservableDecoderRunner(_FakeDecoder(H, truth), L)
    res = run.from_errors_decode_observables_detailed_batch(truth)
    assert res[0].success and not res[0].error_detected and res[0].decoding_quality == truth[0].sum()
    assert not res[1].success and res[1].unconverged_no_error
    for r, e in zip(res[2:], truth[2:]):
        det = bool(((L.astype(int) @ e) % 2).any())
        assert r.error_detected == det and r.unconverged_logical_error == det
    s = run.logical_error_rate(truth, rounds=2)
    assert s["failures"] == 7 and s["unconverged"] == 7
    assert abs(s["per_round_logical_error_rate"] - (1 - (1 - 7 / 8) ** 0.5)) < 1e-12


@pytest.mark.gpu
def test_bb_code_logical_error_rate_on_the_gpu():
    from pelagos_amd import RelayBPDecoder

    H = codes.bb_72_12_6()
    L = codes.logical_operators(H.toarray(), codes.bb_72_12_6_hz())
    p = np.full(72, 0.03)
    dec = RelayBPDecoder(H, p, data_type="float32", algo="relay", seed=0, device=[0])
    run = ObservableDecoderRunner(dec, L)
    err = codes.sample_errors(p, 2048, seed=7)
    print("[gpu] bb72 LER", flush=True)
    s = run.logical_error_rate(err)
    res = run.from_errors_decode_observables_detailed_batch(err[:64])
    # every successful decoding satisfies the syndrome; a decoding that equals the
    # truth predicts the true observables
    for r, e in zip(res, err[:64]):
        if r.success:
            assert ((H @ r.decoding.astype(int)) % 2 == (H @ e.astype(int)) % 2).all()
        if (r.decoding == e).all():
            assert not r.error_detected
    assert s["shots"] == 2048 and 0.0 <= s["logical_error_rate"] < 0.05
