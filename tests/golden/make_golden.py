"""Write the committed fixtures of tests/golden/ (run from the repository root:
``python tests/golden/make_golden.py``).

The reference stores no golden vectors for its decoder (its tests compute
every expected value live from the relay-bp Rust crate, which is not available
here), so the fixtures are of two kinds:

* known answers fixed by the mathematics, computed here by brute force and
  independent of any decoder: the check-node update of exhaustive small rows
  (``prod sign * min over the others`` with the syndrome sign,
  kernels.py:1128-1215) and the maximum-likelihood answers of a chain code,
  where min-sum is exact;
* frozen decodes of the CPU oracle on small seeded workloads ([[72,12,6]] BB
  code, a small banded circuit-shaped code) with the gamma table stored in the
  fixture, so the HIP decoder can be checked against them on a machine that
  does not build the oracle, and a change of the oracle itself is noticed.
"""

from __future__ import annotations

import itertools
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

OUT = Path(__file__).resolve().parent
RELAY = dict(pre_iter=80, num_sets=100, set_max_iter=60, stop_nconv=5, gamma0=0.65)


def check_update_vectors() -> dict:
    """Every row of degree 1-4 over a value set with ties, zeros and both signs."""
    vals = np.float32([-2.0, -0.5, 0.0, 0.5, 1.0, 2.0])
    rows, degs, syns, nus = [], [], [], []
    for d in range(1, 5):
        for mu in itertools.product(vals, repeat=d):
            mu = np.float32(mu)
            for s in (0, 1):
                nu = np.zeros(d, np.float32)
                if d > 1:
                    for e in range(d):
                        others = np.delete(mu, e)
                        sign = (-1) ** s * np.prod(np.where(others < 0, -1.0, 1.0))
                        nu[e] = np.float32(sign * np.min(np.abs(others)))
                row = np.zeros(4, np.float32)
                row[:d] = mu
                out = np.zeros(4, np.float32)
                out[:d] = nu
                rows.append(row)
                degs.append(d)
                syns.append(s)
                nus.append(out)
    return dict(mu=np.array(rows), deg=np.array(degs, np.int32), syn=np.array(syns, np.uint8),
                nu=np.array(nus))


def chain_ml() -> dict:
    """Repetition (chain) code of 9 bits, uniform prior 0.1: for every syndrome,
    the weight of the most likely error pattern, by enumeration of all 512."""
    import toy_codes

    n = 9
    H = toy_codes.repetition_code(n).toarray().astype(np.int64)
    pats = np.array(list(itertools.product([0, 1], repeat=n)), np.int64)
    syn_of = (pats @ H.T) % 2
    llr = np.float32(np.log(0.9 / 0.1))
    best = {}
    for e, s in zip(pats, syn_of):
        key = tuple(s)
        w = int(e.sum())
        if key not in best or w < best[key][0]:
            best[key] = (w, e)
    keys = sorted(best)
    syn = np.array(keys, np.uint8)
    ml_w = np.array([best[k][0] for k in keys], np.int32)
    ml_e = np.array([best[k][1] for k in keys], np.uint8)
    return dict(H=H.astype(np.uint8), syn=syn, ml_weight=ml_w, ml_errors=ml_e, llr=np.float32(llr))


def gamma_rows_from_codes(codes_u8: np.ndarray) -> np.ndarray:
    """Memory strengths on a 1/255 grid of (-0.24, 0.66), stored as uint8 codes
    so the fixture stays small; the float32 values are rebuilt exactly."""
    return (np.float32(-0.24) + np.float32(0.9) * (codes_u8.astype(np.float32) / np.float32(255.0))).astype(
        np.float32)


def frozen_decode(H, p, shots: int, seed: int, num_sets: int) -> dict:
    import oracle
    from pelagos_amd import codes, utils
    from pelagos_amd.graph import CSRGraph

    g = CSRGraph.from_matrix(H)
    err = codes.sample_errors(p, shots, seed=seed)
    syn = codes.syndromes_of(H, err)
    lam = utils.compute_log_prior_ratios(p)
    gcodes = np.random.default_rng(seed + 1000).integers(0, 256, (num_sets, g.V)).astype(np.uint8)
    gtab = gamma_rows_from_codes(gcodes)
    kw = dict(RELAY, num_sets=num_sets)
    ref = oracle.decode(g.chk_ptr, g.edge_var, g.V, syn, lam, gtab, want_posterior=False, nthreads=8, **kw)
    Hc = H.tocsr()
    return dict(indptr=Hc.indptr.astype(np.int32), indices=Hc.indices.astype(np.int32),
                shape=np.array(Hc.shape, np.int32), priors=np.asarray(p, np.float64), syndromes=syn,
                gamma_codes=gcodes, num_sets=np.int32(num_sets), errors=np.packbits(ref["errors"], axis=1),
                weights=ref["weights"], iterations=ref["iterations"], found_count=ref["found_count"])


def main() -> None:
    import oracle
    from pelagos_amd import codes

    oracle.build()
    np.savez_compressed(OUT / "check_update.npz", **check_update_vectors())
    np.savez_compressed(OUT / "chain9_ml.npz", **chain_ml())
    np.savez_compressed(OUT / "bb72_relay.npz",
                        **frozen_decode(codes.bb_72_12_6(), np.full(72, 0.03), 128, seed=21, num_sets=100))
    Hb, pb = codes.banded_circuit_code(n_blocks=4, V=2700, seed=3)
    np.savez_compressed(OUT / "banded4x72_relay.npz", **frozen_decode(Hb, pb, 48, seed=22, num_sets=24))
    for f in sorted(OUT.glob("*.npz")):
        print(f.name, f.stat().st_size)


if __name__ == "__main__":
    main()
