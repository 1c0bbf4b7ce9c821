"""Tanner-graph normalisation on the host.

The reference's ``CSRGraph`` (src/pelagos/graph.py:31-211) numbers edges in
CSR (check-major) order and groups them per variable.  This module performs
the same coercion of the input matrix (graph.py:66-82) and hands the CSR
arrays to the native library, which builds the device layout (check rows,
variable lists in ascending edge id, sign-bit words) in ``pel_graph_create``.
"""

from __future__ import annotations

import hashlib
from dataclasses import dataclass, field

import numpy as np
from scipy.sparse import csr_matrix, issparse


@dataclass
class CSRGraph:
    """Check-major CSR of H plus the variable-major view (ascending edge id).

    Attributes mirror graph.py:156-181: ``chk_ptr [C+1]``, ``edge_var [E]``
    (the column of each edge), ``var_ptr [V+1]``, ``var_edges [E]``,
    ``edge_chk [E]``, degree maxima and the sparsity hash.
    """

    C: int
    V: int
    E: int
    chk_ptr: np.ndarray
    edge_var: np.ndarray
    var_ptr: np.ndarray = field(repr=False)
    var_edges: np.ndarray = field(repr=False)
    edge_chk: np.ndarray = field(repr=False)
    max_deg_chk: int = 0
    max_deg_var: int = 0
    sparsity_hash: str = ""

    @classmethod
    def from_matrix(cls, H) -> "CSRGraph":
        if not issparse(H):
            raise ValueError("Input must be a scipy.sparse matrix")
        if not isinstance(H, csr_matrix):
            H = H.tocsr()
        if H.dtype != np.uint8 and H.dtype != bool:
            H = csr_matrix((H != 0).astype(np.uint8))
        C, V = H.shape
        chk_ptr = np.ascontiguousarray(H.indptr, dtype=np.int32)
        edge_var = np.ascontiguousarray(H.indices, dtype=np.int32)
        E = int(chk_ptr[-1])
        edge_chk = np.repeat(np.arange(C, dtype=np.int32), np.diff(chk_ptr))
        # Stable sort: within a variable, edges keep ascending id (= ascending check).
        var_edges = np.argsort(edge_var, kind="stable").astype(np.int32)
        var_ptr = np.zeros(V + 1, dtype=np.int32)
        np.cumsum(np.bincount(edge_var, minlength=V), out=var_ptr[1:])
        chk_deg = np.diff(chk_ptr)
        var_deg = np.diff(var_ptr)
        blob = np.concatenate([
            np.sort(chk_deg).astype(np.int32),
            np.sort(var_deg).astype(np.int32),
            np.concatenate([np.sort(edge_var[chk_ptr[i]:chk_ptr[i + 1]]) for i in range(C)]
                           ).astype(np.int32) if E else np.zeros(0, np.int32),
        ]).tobytes()
        return cls(
            C=C, V=V, E=E, chk_ptr=chk_ptr, edge_var=edge_var, var_ptr=var_ptr,
            var_edges=var_edges, edge_chk=edge_chk,
            max_deg_chk=int(chk_deg.max()) if C else 0,
            max_deg_var=int(var_deg.max()) if V else 0,
            sparsity_hash=hashlib.blake2b(blob, digest_size=16).hexdigest(),
        )

    def to_dense(self) -> np.ndarray:
        H = np.zeros((self.C, self.V), dtype=np.uint8)
        H[self.edge_chk, self.edge_var] ^= 1
        return H

    def syndrome(self, errors: np.ndarray) -> np.ndarray:
        """H e mod 2 for ``errors [B, V]`` (uint8), computed edge by edge."""
        errors = np.asarray(errors, dtype=np.uint8)
        H = csr_matrix((np.ones(self.E, np.int32), self.edge_var, self.chk_ptr),
                       shape=(self.C, self.V))
        return ((H @ errors.T.astype(np.int32)) % 2).T.astype(np.uint8)

    def get_stats(self) -> dict:
        return {
            "num_checks": self.C, "num_vars": self.V, "num_edges": self.E,
            "max_deg_chk": self.max_deg_chk, "max_deg_var": self.max_deg_var,
            "sparsity_hash": self.sparsity_hash,
        }
