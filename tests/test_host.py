"""CPU tests of the host side: C-ABI library, priors, gamma draws, bit packing,
graph layout and the workload generators (tests/test_cpu_units.py and
tests/test_gamma_sampling.py of the reference are the model)."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest
import torch
from scipy.sparse import csr_matrix

import toy_codes
from pelagos_amd import _native, codes, utils
from pelagos_amd.decoder import alpha_schedule
from pelagos_amd.graph import CSRGraph

ROOT = Path(__file__).resolve().parent.parent


# --- the C ABI ----------------------------------------------------------------------------------


def _header_symbols():
    text = (ROOT / "include" / "pelagos_relaybp.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char \*)\s*(pel_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = _native.lib()
    syms = _header_symbols()
    assert syms == sorted(_native.SIGNATURES), "header and binding disagree"
    for name in syms:
        assert hasattr(L, name), name
    assert L.pel_abi_version() == _native.PEL_ABI_VERSION


def test_library_rejects_bad_graphs_without_a_gpu():
    """Argument validation happens before any device call."""
    L = _native.lib()
    h = ctypes.c_void_p()
    ptr = np.array([0, 2], np.int32)
    var = np.array([0, 5], np.int32)           # column 5 out of range for V=3
    rc = L.pel_graph_create(ptr.ctypes.data, var.ctypes.data, 1, 3, 2, ctypes.byref(h))
    assert rc == _native.PEL_EINVAL
    assert b"out of range" in L.pel_last_error()
    bad_ptr = np.array([0, 3], np.int32)
    rc = L.pel_graph_create(bad_ptr.ctypes.data, var.ctypes.data, 1, 6, 2, ctypes.byref(h))
    assert rc == _native.PEL_EINVAL
    assert L.pel_graph_destroy(None) == 0
    assert L.pel_plan_decode(None, 4, 0, -1, 0, None) == _native.PEL_EINVAL
    assert L.pel_decode(None, None, None, None, None) == _native.PEL_EINVAL


# --- priors (tests/test_cpu_units.py:45-65) -------------------------------------------------------


def test_log_prior_ratios_match_formula():
    p = np.array([0.001, 0.01, 0.1, 0.3])
    llr = utils.compute_log_prior_ratios(p)
    assert llr.dtype == np.float32
    np.testing.assert_allclose(llr, np.log((1 - p) / p), rtol=1e-6)


def test_log_prior_ratios_are_finite_for_zero_prior():
    llr = utils.compute_log_prior_ratios(np.array([0.0, 0.05]))
    assert np.all(np.isfinite(llr))
    assert llr[0] > llr[1]
    assert llr[0] == np.float32(np.log((1 - 1e-10) / 1e-10))


def test_validate_error_priors():
    p = np.array([0.001, 0.2, 0.45])
    np.testing.assert_array_equal(utils.validate_error_priors(p), p)
    with pytest.raises(ValueError, match="1D"):
        utils.validate_error_priors(np.zeros((2, 2)))


def test_validate_csr_matrix():
    with pytest.raises(ValueError, match="scipy.sparse"):
        utils.validate_csr_matrix(np.eye(3))
    with pytest.raises(ValueError, match="empty"):
        utils.validate_csr_matrix(csr_matrix((0, 3)))
    with pytest.raises(ValueError, match="zero non-zero"):
        utils.validate_csr_matrix(csr_matrix((2, 3)))
    assert utils.validate_csr_matrix(toy_codes.triangle_code().tocoo())


# --- bit packing (tests/test_cpu_units.py:93-99) ----------------------------------------------------


@pytest.mark.parametrize("V", [1, 31, 32, 33, 100])
def test_bitpack_round_trip(V):
    e = torch.from_numpy((np.random.default_rng(V).random((7, V)) < 0.5).astype(np.uint8))
    packed = utils.bitpack_errors(e)
    assert packed.dtype == torch.int32 and packed.shape == (7, (V + 31) // 32)
    assert torch.equal(utils.bitunpack_errors(packed, V), e)
    # bit i of word w is variable 32 w + i
    one = torch.zeros(1, V, dtype=torch.uint8)
    one[0, V - 1] = 1
    w = utils.bitpack_errors(one)[0, (V - 1) // 32].item() & 0xFFFFFFFF
    assert w == 1 << ((V - 1) % 32)


# --- gamma draws (tests/test_gamma_sampling.py:49-103) -----------------------------------------------


def test_seeded_gamma_equals_manual_seed_rand():
    g = utils.sample_gamma_uniform(2, 50, -0.24, 0.66, "cpu", seed=11)
    torch.manual_seed(11)
    u = torch.rand(2, 50)
    assert torch.equal(g, -0.24 + (0.66 - -0.24) * u)


def test_gamma_draw_leaves_global_rng_alone():
    torch.manual_seed(5)
    a = torch.rand(3)
    torch.manual_seed(5)
    utils.sample_gamma_uniform(1, 10, 0.0, 1.0, "cpu", seed=99)
    b = torch.rand(3)
    assert torch.equal(a, b)


def test_concurrent_gamma_draws_match_sequential():
    """Seeded draws from several threads at once equal the same draws made one
    after another: each call has its own generator (the reference's
    test_gamma_sampling.py checks the same of its samplers)."""
    from concurrent.futures import ThreadPoolExecutor

    intervals = torch.tensor([[-0.24, 0.66], [0.0, 0.5], [0.1, 0.2]] * 20)
    jobs = [(seed, kind) for seed in range(24) for kind in ("uniform", "per_variable")]

    def draw(job):
        seed, kind = job
        if kind == "uniform":
            return utils.sample_gamma_uniform(4, 60, -0.24, 0.66, "cpu", seed=seed)
        return utils.sample_gamma_per_variable(4, 60, intervals, "cpu", seed=seed)

    sequential = [draw(j) for j in jobs]
    with ThreadPoolExecutor(max_workers=8) as pool:
        concurrent = list(pool.map(draw, jobs))
    for a, b in zip(sequential, concurrent):
        assert torch.equal(a, b)


def test_per_variable_gamma_respects_intervals():
    V = 40
    iv = torch.stack([torch.linspace(-0.5, 0.0, V), torch.linspace(0.1, 0.9, V)], 1)
    g = utils.sample_gamma_per_variable(64, V, iv, "cpu", seed=3)
    assert (g >= iv[:, 0]).all() and (g <= iv[:, 1]).all()
    with pytest.raises(ValueError, match="gamma_intervals"):
        utils.sample_gamma_per_variable(1, V + 1, iv, "cpu", seed=3)


def test_explicit_gamma_rows_rotate_by_one():
    """The reference's known answer (tests/test_decoder_fixes.py:400-429): relay
    leg L starts with explicit_gammas[(L + 1) % K]; for K = 3 over 7 legs that is
    rows 1, 2, 0, 1, 2, 0, 1, and a single row serves every leg."""
    assert utils.explicit_leg_rows(3, 7) == [1, 2, 0, 1, 2, 0, 1]
    assert utils.explicit_leg_rows(1, 4) == [0, 0, 0, 0]
    assert utils.explicit_leg_rows(5, 0) == []
    with pytest.raises(ValueError):
        utils.explicit_leg_rows(0, 2)


def test_alpha_schedule():
    a = alpha_schedule(0.0, 2.0, 4)
    np.testing.assert_array_equal(a, np.float32([1 - 2 ** -(t / 2) for t in range(1, 5)]))
    np.testing.assert_array_equal(alpha_schedule(None, 1.0, 3), np.ones(3, np.float32))
    np.testing.assert_array_equal(alpha_schedule(0.75, 1.0, 2), np.float32([0.75, 0.75]))


# --- graph and generators ----------------------------------------------------------------------------


def test_graph_orders_edges_like_the_reference():
    H = csr_matrix(np.array([[1, 0, 1, 1], [0, 1, 1, 0], [1, 1, 0, 1]], np.uint8))
    g = CSRGraph.from_matrix(H)
    assert g.E == 8
    np.testing.assert_array_equal(g.chk_ptr, [0, 3, 5, 8])
    np.testing.assert_array_equal(g.edge_var, [0, 2, 3, 1, 2, 0, 1, 3])
    # per variable, ascending edge id
    np.testing.assert_array_equal(g.var_ptr, [0, 2, 4, 6, 8])
    np.testing.assert_array_equal(g.var_edges, [0, 5, 3, 6, 1, 4, 2, 7])
    np.testing.assert_array_equal(g.to_dense(), H.toarray())
    e = np.array([[1, 0, 0, 1], [0, 1, 1, 0]], np.uint8)
    np.testing.assert_array_equal(g.syndrome(e), (e @ H.toarray().T) % 2)


def test_graph_coerces_non_binary_values():
    H = csr_matrix(np.array([[2, 0, 1], [0, 3, 0]], np.int64))
    g = CSRGraph.from_matrix(H)
    assert g.E == 3 and g.to_dense().max() == 1


def test_bb_code_properties():
    hx = codes.bb_72_12_6().toarray()
    hz = codes.bb_72_12_6_hz()
    assert hx.shape == (36, 72)
    assert set(hx.sum(1)) == {6} and set(hx.sum(0)) == {3}
    assert codes.gf2_rank(hx) == 30
    assert not ((hx.astype(int) @ hz.T) % 2).any()
    L = codes.logical_operators(hx, hz)
    assert L.shape == (12, 72)
    assert not ((L.astype(int) @ hz.T) % 2).any()


def test_banded_code_shape():
    H, p = codes.banded_circuit_code(seed=0)
    assert H.shape == (936, 8785)
    deg = np.diff(H.tocsc().indptr)
    assert deg.min() >= 3 and deg.max() <= 6
    rows, cols = H.nonzero()
    blk = rows // 72
    lo = np.full(8785, 99)
    np.minimum.at(lo, cols, blk)
    hi = np.zeros(8785, int)
    np.maximum.at(hi, cols, blk)
    assert (hi - lo <= 1).all()
    assert p.dtype == np.float32 and p.min() >= 1e-4 and p.max() <= 1e-2
    H2, _ = codes.banded_circuit_code(n_blocks=26, V=17570, seed=0)
    assert H2.shape == (1872, 17570)
