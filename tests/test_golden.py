"""Committed fixtures (tests/golden/, written by tests/golden/make_golden.py).

* check_update.npz: the check-node update of every row of degree 1-4 over a
  value set with ties, zeros and both signs, computed by brute force from its
  definition (kernels.py:1128-1215: sign product with the syndrome sign times
  the minimum over the other edges; degree-1 checks send 0).
* chain9_ml.npz: maximum-likelihood answers of the 9-bit chain code for all
  256 syndromes, by enumeration; min-sum on a tree Tanner graph is exact.
* bb72_relay.npz, banded4x72_relay.npz: frozen decodes of the CPU oracle with
  the full Relay-BP-S schedule; the gamma rows are stored with the fixture.

The CPU tests check the oracle against all four; the GPU tests check the HIP
decoder, through the C ABI, against the same files.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest
import torch
from scipy.sparse import csr_matrix

import oracle

GOLD = Path(__file__).resolve().parent / "golden"
RELAY = dict(pre_iter=80, set_max_iter=60, stop_nconv=5, gamma0=0.65)
FROZEN = ["bb72_relay.npz", "banded4x72_relay.npz"]


def _gamma_rows(codes_u8: np.ndarray) -> np.ndarray:
    return (np.float32(-0.24) + np.float32(0.9) * (codes_u8.astype(np.float32) / np.float32(255.0))).astype(
        np.float32)


def _load(name):
    return dict(np.load(GOLD / name, allow_pickle=False))


def _matrix(f):
    return csr_matrix((np.ones(f["indices"].shape[0], np.uint8), f["indices"], f["indptr"]),
                      shape=tuple(int(x) for x in f["shape"]))


# --- CPU: the oracle against the fixtures --------------------------------------------------------


def test_oracle_check_update_matches_known_answers():
    f = _load("check_update.npz")
    for mu, d, s, nu in zip(f["mu"], f["deg"], f["syn"], f["nu"]):
        got = oracle.check_update_row(mu[:d], int(s))
        np.testing.assert_array_equal(got, nu[:d], err_msg=f"mu={mu[:d]} s={s}")


def test_oracle_chain_code_is_maximum_likelihood():
    f = _load("chain9_ml.npz")
    H = csr_matrix(f["H"])
    from pelagos_amd.graph import CSRGraph

    g = CSRGraph.from_matrix(H)
    lam = np.full(H.shape[1], f["llr"], np.float32)
    out = oracle.decode(g.chk_ptr, g.edge_var, g.V, f["syn"], lam, np.zeros((1, g.V), np.float32),
                        pre_iter=20, num_sets=0, set_max_iter=1, stop_nconv=1, gamma0=0.0, plain=True)
    assert (out["found_count"] == 1).all()
    np.testing.assert_array_equal(np.rint(out["weights"] / f["llr"]).astype(np.int32), f["ml_weight"])


@pytest.mark.parametrize("name", FROZEN)
def test_oracle_reproduces_frozen_decodes(name):
    f = _load(name)
    H = _matrix(f)
    from pelagos_amd import utils
    from pelagos_amd.graph import CSRGraph

    g = CSRGraph.from_matrix(H)
    lam = utils.compute_log_prior_ratios(f["priors"])
    out = oracle.decode(g.chk_ptr, g.edge_var, g.V, f["syndromes"], lam, _gamma_rows(f["gamma_codes"]),
                        num_sets=int(f["num_sets"]), nthreads=4, **RELAY)
    np.testing.assert_array_equal(np.packbits(out["errors"], axis=1), f["errors"])
    for k in ("weights", "iterations", "found_count"):
        np.testing.assert_array_equal(out[k], f[k], err_msg=k)


# --- GPU: the HIP decoder against the same fixtures ---------------------------------------------


@pytest.mark.gpu
@pytest.mark.parametrize("name", FROZEN)
def test_gpu_reproduces_frozen_decodes(name):
    from pelagos_amd import RelayBPDecoder

    f = _load(name)
    H = _matrix(f)
    rows = _gamma_rows(f["gamma_codes"])
    # explicit_gammas: leg l uses row (l + 1) % K (decoder.py:3685-3688)
    dec = RelayBPDecoder(H, f["priors"], num_sets=int(f["num_sets"]), gamma_dist_interval=(-0.24, 0.66),
                         data_type="float32", algo="relay", seed=0, device=[0],
                         explicit_gammas=np.roll(rows, 1, axis=0), **RELAY)
    print(f"[gpu] golden {name} B={f['syndromes'].shape[0]}", flush=True)
    out = dec.decode(torch.from_numpy(f["syndromes"]).to("cuda:0"))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(np.packbits(out["errors"].cpu().numpy(), axis=1), f["errors"])
    for k in ("weights", "iterations", "found_count"):
        np.testing.assert_array_equal(out[k].cpu().numpy(), f[k], err_msg=k)


@pytest.mark.gpu
def test_gpu_chain_code_is_maximum_likelihood():
    from pelagos_amd import RelayBPDecoder

    f = _load("chain9_ml.npz")
    H = csr_matrix(f["H"])
    dec = RelayBPDecoder(H, np.full(H.shape[1], 0.1), pre_iter=20, num_sets=0, set_max_iter=1, stop_nconv=1,
                         gamma0=0.0, data_type="float32", algo="plain", seed=0, device=[0])
    print("[gpu] golden chain9", flush=True)
    out = dec.decode(torch.from_numpy(f["syn"]).to("cuda:0"))
    torch.cuda.synchronize()
    w = out["weights"].cpu().numpy()
    np.testing.assert_array_equal(np.rint(w / f["llr"]).astype(np.int32), f["ml_weight"])
