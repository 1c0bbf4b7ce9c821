"""Host-side helpers of the Relay-BP-S path: priors, memory-strength draws, bit packing.

Each function keeps the reference's name, arguments and arithmetic
(src/pelagos/utils.py:32-272) so a caller can switch imports.  The memory
strength draws issue exactly the reference's torch calls on the decoding
device, which is what makes the per-leg gamma rows identical.
"""

from __future__ import annotations

from typing import Optional

import numpy as np
import torch
from scipy.sparse import csr_matrix, issparse

GAMMA_DRAW_MODES = ("broadcast", "per_row")
UNCONVERGED_CORRECTIONS = ("first_leg", "last_leg")


def compute_log_prior_ratios(error_priors: np.ndarray) -> np.ndarray:
    """ln((1-p)/p) after clipping p to [1e-10, 0.5-1e-10], computed in float64
    and returned as float32 (utils.py:88-107)."""
    p = np.clip(np.asarray(error_priors, dtype=np.float64), 1e-10, 0.5 - 1e-10)
    return np.log((1 - p) / p).astype(np.float32)


def validate_error_priors(error_priors: np.ndarray) -> np.ndarray:
    """1-D check and clip to [1e-10, 0.5-1e-10] (utils.py:231-251)."""
    error_priors = np.asarray(error_priors)
    if error_priors.ndim != 1:
        raise ValueError("error_priors must be 1D array")
    clipped = np.clip(error_priors, 1e-10, 0.5 - 1e-10)
    if np.any(clipped != error_priors):
        print("Warning: error_priors were clipped to avoid numerical issues")
    return clipped


def validate_csr_matrix(H_csr) -> bool:
    """Reject non-sparse, empty or all-zero check matrices (utils.py:254-272)."""
    if not issparse(H_csr):
        raise ValueError("Input must be a scipy.sparse matrix (any format)")
    if not isinstance(H_csr, csr_matrix):
        H_csr = H_csr.tocsr()
    if H_csr.shape[0] == 0 or H_csr.shape[1] == 0:
        raise ValueError("Matrix cannot be empty")
    if H_csr.nnz == 0:
        raise ValueError("Matrix cannot have zero non-zero elements")
    if not np.all((H_csr.data == 0) | (H_csr.data == 1)):
        print("Warning: Matrix contains non-binary values, will be treated as binary")
    return True


def sample_gamma_uniform(rows: int, V: int, gamma_min: float, gamma_max: float, device,
                         seed: Optional[int] = None) -> torch.Tensor:
    """[rows, V] draw of U(gamma_min, gamma_max) from a generator private to the
    call, seeded ``seed`` (utils.py:128-160).  The global RNG is untouched."""
    gen = None
    if seed is not None:
        gen = torch.Generator(device=device)
        gen.manual_seed(seed)
    uniform = torch.rand(rows, V, device=device, generator=gen)
    return gamma_min + (gamma_max - gamma_min) * uniform


def sample_gamma_scalar(B: int, V: int, gamma_value: float, device) -> torch.Tensor:
    """[B, V] filled with gamma_value (utils.py:163-178)."""
    return torch.full((B, V), gamma_value, device=device)


def sample_gamma_per_variable(rows: int, V: int, gamma_intervals: torch.Tensor, device,
                              seed: Optional[int] = None) -> torch.Tensor:
    """[rows, V] draw of U(low_v, high_v) per variable (utils.py:181-225)."""
    gen = None
    if seed is not None:
        gen = torch.Generator(device=device)
        gen.manual_seed(seed)
    if gamma_intervals.ndim != 2 or gamma_intervals.shape[0] != V or gamma_intervals.shape[1] != 2:
        raise ValueError(
            f"gamma_intervals must have shape (V={V}, 2), got {tuple(gamma_intervals.shape)}"
        )
    uniform = torch.rand(rows, V, device=device, generator=gen)
    low = gamma_intervals[:, 0].unsqueeze(0)
    high = gamma_intervals[:, 1].unsqueeze(0)
    return low + (high - low) * uniform


def explicit_leg_rows(K: int, R: int) -> list:
    """Row of an explicit [K, V] gamma table that relay leg l = 0..R-1 uses:
    (l + 1) % K, the Rust decoder's rotation the reference reproduces
    (decoder.py:3685-3688)."""
    if K <= 0:
        raise ValueError("explicit_gammas must have at least one row")
    return [(leg + 1) % K for leg in range(R)]


def bitpack_errors(errors: torch.Tensor, bits_per_word: int = 32) -> torch.Tensor:
    """[B, V] 0/1 -> [B, ceil(V/32)] int32, bit i of word w = variable 32*w + i
    (utils.py:32-60)."""
    B, V = errors.shape
    words = (V + bits_per_word - 1) // bits_per_word
    padded = torch.zeros(B, words * bits_per_word, dtype=torch.int32, device=errors.device)
    padded[:, :V] = errors.to(torch.int32)
    shifts = torch.arange(bits_per_word, dtype=torch.int32, device=errors.device)
    packed = (padded.view(B, words, bits_per_word) << shifts).sum(dim=2, dtype=torch.int64)
    # Bit 31 set means a negative int32: wrap from int64 like the reference's in-place |=.
    packed = torch.where(packed >= 2**31, packed - 2**32, packed)
    return packed.to(torch.int32)


def bitunpack_errors(packed: torch.Tensor, V: int, bits_per_word: int = 32) -> torch.Tensor:
    """Inverse of :func:`bitpack_errors` (utils.py:63-85)."""
    B, words = packed.shape
    shifts = torch.arange(bits_per_word, dtype=torch.int32, device=packed.device)
    bits = ((packed.unsqueeze(2) >> shifts) & 1).to(torch.uint8)
    return bits.reshape(B, words * bits_per_word)[:, :V]
