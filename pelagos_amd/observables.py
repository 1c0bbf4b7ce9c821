"""Observable / logical-error-rate harness around the MI355X decoder.

``ObservableDecoderRunner`` keeps the reference runner's surface
(src/pelagos/adapter.py:515-858): from a batch of true error patterns it forms
the syndromes, decodes them, predicts the observables ``O e mod 2`` and
returns one ``DecodeResult`` per shot with the reference's fields
(adapter.py:40-82): the decoding, success (at least one valid solution), the
iteration count, the decoding quality (sum of the selection weights of the
decoded bits), and the four error classes the benchmark counts.
``logical_error_rate`` adds the benchmark's convention
(pelagos_benchmark.py:1666-1672): a shot fails when its observables differ or
it did not converge, and the per-round rate is 1 - (1 - LER)^(1/rounds).
Intervals are Wilson 95% intervals; the reference's +-1 sigma
(sqrt(LER (1 - LER) / N)) is reported beside them.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, List, Optional

import numpy as np
import torch
from scipy import sparse as sp


@dataclass
class DecodeResult:
    decoding: np.ndarray
    success: bool
    iterations: int
    decoding_quality: float = float("inf")
    unconverged_no_error: Optional[bool] = None
    unconverged_logical_error: Optional[bool] = None
    better_decoding_quality_error: Optional[bool] = None
    worse_decoding_quality_error: Optional[bool] = None
    converged: bool = False
    error_detected: bool = False
    observables: Optional[np.ndarray] = None
    unconverged_indices: Optional[np.ndarray] = None
    found_count: Optional[int] = None


def wilson(k: int, n: int, z: float = 1.959964):
    if n == 0:
        return (0.0, 0.0, 1.0)
    p = k / n
    den = 1 + z * z / n
    c = (p + z * z / (2 * n)) / den
    h = z * math.sqrt(p * (1 - p) / n + z * z / (4 * n * n)) / den
    return (p, max(0.0, c - h), min(1.0, c + h))


def _mul_mod2(A, x: np.ndarray) -> np.ndarray:
    """(A @ x) mod 2 for a 0/1 matrix A (sparse or dense) and 0/1 columns x."""
    if sp.issparse(A):
        y = A @ np.asarray(x, dtype=np.int32)
    else:
        y = np.asarray(A, dtype=np.int32) @ np.asarray(x, dtype=np.int32)
    return (np.asarray(y) % 2).astype(np.uint8)


class ObservableDecoderRunner:
    """decoder: a RelayBPDecoder; observables_matrix: [O, V] 0/1 (dense or sparse)."""

    def __init__(self, decoder, observables_matrix, include_decode_result: bool = True, **kwargs):
        self.decoder = decoder
        self._obs = sp.csr_matrix(observables_matrix) if not sp.issparse(observables_matrix) \
            else observables_matrix.tocsr()
        self.O, self.N = self._obs.shape
        if self.N != decoder.V:
            raise ValueError(f"observables have {self.N} columns, the decoder {decoder.V} variables")
        self.M = int(decoder.C)
        self.include_decode_result = include_decode_result

    def _decode(self, errors: np.ndarray) -> Dict[str, np.ndarray]:
        errors = np.asarray(errors, dtype=np.uint8)
        if errors.ndim != 2 or errors.shape[1] != self.N:
            raise ValueError(f"errors must be [B, {self.N}]")
        H = self.decoder.H_csr
        syn = _mul_mod2(H, errors.T).T
        out = self.decoder.decode(torch.from_numpy(np.ascontiguousarray(syn)).to(self.decoder.device))
        bits = out["errors"]
        if self.decoder.bitpack_output:
            from .utils import bitunpack_errors

            bits = bitunpack_errors(bits, self.N)
        return dict(
            bits=bits.to(torch.uint8).cpu().numpy(),
            iterations=out["iterations"].cpu().numpy(),
            found=out["found_count"].cpu().numpy(),
            unconverged=out["unconverged_indices"].cpu().numpy(),
            true_obs=_mul_mod2(self._obs, errors.T).T,
            errors=errors,
        )

    def from_errors_decode_observables_detailed_batch(self, errors: np.ndarray,
                                                      parallel: bool = False) -> List[DecodeResult]:
        r = self._decode(errors)
        bits, found = r["bits"], r["found"]
        pred = _mul_mod2(self._obs, bits.T).T
        wj = self.decoder.wj.cpu().numpy().astype(np.float64)
        dq = bits.astype(np.float64) @ wj
        tq = r["errors"].astype(np.float64) @ wj
        res = []
        for i in range(bits.shape[0]):
            det = bool((pred[i] != r["true_obs"][i]).any())
            ok = bool(found[i] > 0)
            q = float(dq[i]) if ok else float("inf")
            res.append(DecodeResult(
                decoding=bits[i], success=ok, iterations=int(r["iterations"][i]), decoding_quality=q,
                unconverged_no_error=bool(not det and not ok), unconverged_logical_error=bool(det and not ok),
                better_decoding_quality_error=bool(ok and det and q < tq[i]),
                worse_decoding_quality_error=bool(ok and det and not q < tq[i]),
                converged=ok, error_detected=det, observables=pred[i], found_count=int(found[i])))
        return res

    def logical_error_rate(self, errors: np.ndarray, rounds: int = 1) -> Dict[str, object]:
        """Failures = observables differ, or no solution (converged = found >= 1,
        adapter.py:676-678); summary over the batch."""
        res = self.from_errors_decode_observables_detailed_batch(errors)
        n = len(res)
        det = sum(r.error_detected and r.success for r in res)
        unconv = sum(not r.success for r in res)
        k = det + unconv
        ler, lo, hi = wilson(k, n)
        return {
            "shots": n, "failures": k, "observable_errors": det, "unconverged": unconv,
            "logical_error_rate": ler, "wilson95": [lo, hi],
            "one_sigma": math.sqrt(ler * (1 - ler) / n) if n else 0.0,
            "per_round_logical_error_rate": 1 - (1 - ler) ** (1.0 / max(1, rounds)),
            "mean_iterations": float(np.mean([r.iterations for r in res])) if n else 0.0,
        }
