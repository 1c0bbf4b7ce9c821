"""Toy check matrices of the reference's tests, used only by this test suite:
(3,6)-regular codes, random column-weight LDPC codes, the triangle, chain
(repetition) and high-degree codes (tests/test_decoder_fixes.py:73-97,
215-220; tests/test_triton_against_rust.py:322, 663-690)."""

from __future__ import annotations

import numpy as np
from scipy.sparse import csr_matrix


def regular_code(seed: int, V: int = 600, C: int = 300) -> csr_matrix:
    """(3,6)-regular check matrix (tests/test_decoder_fixes.py:73-86)."""
    rng = np.random.default_rng(seed)
    while True:
        sockets = np.repeat(np.arange(C), 6)
        rng.shuffle(sockets)
        rows = sockets.reshape(V, 3)
        if all(len(set(r)) == 3 for r in rows):
            break
    return csr_matrix(
        (np.ones(3 * V, np.uint8), (rows.ravel(), np.repeat(np.arange(V), 3))), shape=(C, V)
    )


def random_ldpc(C: int, V: int, col_weight: int, seed: int) -> csr_matrix:
    """Random column-weight LDPC (tests/test_decoder_fixes.py:215-220)."""
    rng = np.random.default_rng(seed)
    H = np.zeros((C, V), dtype=np.uint8)
    for v in range(V):
        H[rng.choice(C, col_weight, replace=False), v] = 1
    return csr_matrix(H)


def triangle_code() -> csr_matrix:
    """3 variables, 2 checks (tests/test_triton_against_rust.py:322)."""
    return csr_matrix(np.array([[1, 1, 0], [0, 1, 1]], dtype=np.uint8))


def repetition_code(n: int) -> csr_matrix:
    """Chain code: check i joins variables i and i+1 (a tree Tanner graph)."""
    H = np.zeros((n - 1, n), dtype=np.uint8)
    for i in range(n - 1):
        H[i, i] = H[i, i + 1] = 1
    return csr_matrix(H)


def high_degree_code() -> csr_matrix:
    """Checks of degree 20, 10 and 3 over 25 variables (test_triton_against_rust.py:663-690)."""
    H = np.zeros((3, 25), dtype=np.uint8)
    H[0, :20] = 1
    H[1, 5:15] = 1
    H[2, [0, 10, 20]] = 1
    return csr_matrix(H)
