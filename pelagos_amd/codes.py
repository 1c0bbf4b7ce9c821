"""Check matrices and shot samplers for the decoding workloads.

* ``bb_72_12_6`` -- the [[72,12,6]] bivariate bicycle code in code capacity:
  H_X = [A | B] with l = m = 6, A = x^3 + y + y^2, B = y^3 + x + x^2 over the
  cyclic shifts x = S_6 (x) I_6 and y = I_6 (x) S_6 (36 x 72, rows of weight 6).
  ``bb_72_12_6_hz`` gives H_Z = [B^T | A^T], whose row space is the set of
  harmless residuals (its logical operators are the checks of the LER).
* ``banded_circuit_code`` -- a synthetic circuit-level-shaped matrix: round
  blocks of 72 checks; every column has weight uniform in {3..6}, its checks
  drawn without replacement from two consecutive blocks; priors log-uniform in
  [1e-4, 1e-2].  936 x 8785 with 13 blocks, 1872 x 17570 with 26 blocks.
* ``three_code_workload`` -- the three codes decoded together by the batched
  and serving workloads, at a given mean prior.
"""

from __future__ import annotations

from typing import Dict, Tuple

import numpy as np
from scipy.sparse import csr_matrix


def _shift(n: int) -> np.ndarray:
    return np.roll(np.eye(n, dtype=np.int64), 1, axis=1)


def _bb_blocks(l: int = 6, m: int = 6) -> Tuple[np.ndarray, np.ndarray]:
    x = np.kron(_shift(l), np.eye(m, dtype=np.int64))
    y = np.kron(np.eye(l, dtype=np.int64), _shift(m))
    mp = np.linalg.matrix_power
    A = (mp(x, 3) + y + mp(y, 2)) % 2
    B = (mp(y, 3) + x + mp(x, 2)) % 2
    return A, B


def bb_72_12_6() -> csr_matrix:
    """H_X = [A | B] of the [[72,12,6]] bivariate bicycle code (36 x 72)."""
    A, B = _bb_blocks()
    return csr_matrix(np.hstack([A, B]).astype(np.uint8))


def bb_72_12_6_hz() -> np.ndarray:
    """H_Z = [B^T | A^T] (dense 36 x 72); H_X H_Z^T = 0 (mod 2)."""
    A, B = _bb_blocks()
    return np.hstack([B.T, A.T]).astype(np.uint8)


def gf2_rank(M: np.ndarray) -> int:
    M = (np.array(M, dtype=np.uint8) & 1).copy()
    r = 0
    rows, cols = M.shape
    for c in range(cols):
        piv = np.nonzero(M[r:, c])[0]
        if piv.size == 0:
            continue
        p = r + piv[0]
        M[[r, p]] = M[[p, r]]
        mask = M[:, c].astype(bool)
        mask[r] = False
        M[mask] ^= M[r]
        r += 1
        if r == rows:
            break
    return r


def gf2_nullspace(M: np.ndarray) -> np.ndarray:
    """Basis (rows) of {v : M v = 0 (mod 2)}."""
    A = (np.array(M, dtype=np.uint8) & 1).copy()
    rows, n = A.shape
    pivots = []
    r = 0
    for c in range(n):
        if r == rows:
            break
        piv = np.nonzero(A[r:, c])[0]
        if piv.size == 0:
            continue
        p = r + piv[0]
        A[[r, p]] = A[[p, r]]
        mask = A[:, c].astype(bool)
        mask[r] = False
        A[mask] ^= A[r]
        pivots.append(c)
        r += 1
    pivset = set(pivots)
    free = [c for c in range(n) if c not in pivset]
    basis = np.zeros((len(free), n), dtype=np.uint8)
    for k, f in enumerate(free):
        basis[k, f] = 1
        for i, pc in enumerate(pivots):
            basis[k, pc] = A[i, f]
    return basis


def logical_operators(hx: np.ndarray, hz: np.ndarray) -> np.ndarray:
    """Rows L (k x n): a basis of ker(H_Z) modulo rowspace(H_X).

    For a residual r = e xor e_hat with H_X r = 0, r lies in rowspace(H_Z)
    (a harmless stabilizer) iff L r = 0 (mod 2) for every row of L.
    """
    base = np.array(hx, dtype=np.uint8) & 1
    rank = gf2_rank(base)
    cur = base
    out = []
    for v in gf2_nullspace(hz):
        trial = np.vstack([cur, v])
        if gf2_rank(trial) > rank:
            cur, rank = trial, rank + 1
            out.append(v)
    return np.array(out, dtype=np.uint8)


def banded_circuit_code(n_blocks: int = 13, checks_per_block: int = 72, V: int = 8785,
                        wmin: int = 3, wmax: int = 6, seed: int = 0
                        ) -> Tuple[csr_matrix, np.ndarray]:
    """Synthetic banded check matrix and log-uniform priors.

    Column j: weight w ~ U{wmin..wmax}, start block b ~ U{0..n_blocks-2}, and w
    distinct checks among the 2*checks_per_block checks of blocks b, b+1.
    Priors p_j = 10**U(-4, -2) as float32.  Returns (H [C x V] uint8 CSR, p).
    """
    rng = np.random.default_rng(seed)
    span = 2 * checks_per_block
    w = rng.integers(wmin, wmax + 1, size=V)
    b0 = rng.integers(0, n_blocks - 1, size=V)
    picks = np.argsort(rng.random((V, span)), axis=1)[:, :wmax]
    rows = b0[:, None] * checks_per_block + picks
    keep = np.arange(wmax)[None, :] < w[:, None]
    r = rows[keep]
    c = np.broadcast_to(np.arange(V)[:, None], rows.shape)[keep]
    C = n_blocks * checks_per_block
    H = csr_matrix((np.ones(r.size, np.uint8), (r, c)), shape=(C, V))
    H.sort_indices()
    p = (10.0 ** rng.uniform(-4.0, -2.0, size=V)).astype(np.float32)
    return H, p


def sample_errors(priors: np.ndarray, shots: int, seed: int = 0, scale: float = 1.0) -> np.ndarray:
    """e_j ~ Bernoulli(scale * p_j), [shots, V] uint8."""
    rng = np.random.default_rng(seed)
    p = np.asarray(priors, dtype=np.float64) * scale
    return (rng.random((shots, p.shape[0])) < p[None, :]).astype(np.uint8)


def syndromes_of(H: csr_matrix, errors: np.ndarray) -> np.ndarray:
    """s = H e mod 2, [shots, C] uint8 (C order, so its buffer can go to the C ABI as is)."""
    return np.ascontiguousarray(((H.astype(np.int32) @ errors.T.astype(np.int32)) % 2).T, dtype=np.uint8)


def synthetic_observables(V: int, k: int = 12, weight: int = 64, seed: int = 0) -> np.ndarray:
    """k seeded random observables over V variables, each the parity of ``weight``
    distinct columns ([k, V] uint8).  A decode of a matrix without a logical
    operator set counts a logical error when O (e xor e_hat) != 0."""
    rng = np.random.default_rng(seed)
    O = np.zeros((k, V), dtype=np.uint8)
    for r in range(k):
        O[r, rng.choice(V, min(weight, V), replace=False)] = 1
    return O


# The three codes of the batched and serving workloads, by name: the [[72,12,6]]
# BB code, the 936 x 8785 banded matrix (13 round-blocks) and a 1872 x 17570
# banded matrix of the same construction over 26 round-blocks.
THREE_CODES = ("bb72", "banded936", "banded1872")


def three_code_workload(prior_scale: float) -> Dict[str, Tuple[csr_matrix, np.ndarray]]:
    """name -> (H, priors) of the three codes with their priors scaled to mean
    ``prior_scale``: uniform for the BB code, the banded matrices' log-uniform
    priors rescaled (capped at 0.45)."""
    out = {"bb72": (bb_72_12_6(), np.full(72, prior_scale, dtype=np.float32))}
    for name, nb, V in (("banded936", 13, 8785), ("banded1872", 26, 17570)):
        H, p = banded_circuit_code(n_blocks=nb, V=V, seed=0)
        out[name] = (H, np.minimum(p * (prior_scale / float(p.mean())), 0.45).astype(np.float32))
    return out
