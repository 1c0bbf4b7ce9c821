"""CPU restatement of the learned memory-strength MLP (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may
import this module, as the checker of the HIP kernels in
pelagos_amd/csrc/gamma_mlp.hip and of the bf16-MFMA inference inside the
decode kernel.  The reference has no learned component; this restates the
network defined in include/pelagos_relaybp.h:

    h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), z = w3 . h2 + b3,
    shift = (hi - lo) * (sigmoid(z) - 1/2)   (added to a learned leg's uniform
                                             draw, then clamped to [lo, hi]),
    loss  = mean_i w_i * (softplus(z_i) - y_i z_i).

Everything runs in float64 so the GPU's fp32 sums can be compared against it
with a stated tolerance.  Its gradient is pinned by central finite
differences in tests/test_learned_gamma.py, and ``shift_q8`` (the kernel's
bf16 inference) by ``forward`` on a network whose every rounding is exact.  At
inference the network runs at the start of relay leg 0, where feature 7 (the
leg index) is 0: the kernel gives that feature's K slot to the bias.
"""

from __future__ import annotations

from typing import Dict, Tuple

import numpy as np

F, H = 8, 32
P = H * F + H + H * H + H + H + 1


def unpack(p: np.ndarray) -> Dict[str, np.ndarray]:
    p = np.asarray(p, np.float64)
    o = 0
    out = {}
    for name, shape in (("W1", (H, F)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("w3", (H,)), ("b3", (1,))):
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    assert o == P
    return out


def forward(p: np.ndarray, x: np.ndarray) -> Tuple[np.ndarray, Dict[str, np.ndarray]]:
    q = unpack(p)
    x = np.asarray(x, np.float64)
    a1 = x @ q["W1"].T + q["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ q["W2"].T + q["b2"]
    h2 = np.maximum(a2, 0.0)
    z = h2 @ q["w3"] + q["b3"][0]
    return z, dict(x=x, a1=a1, h1=h1, a2=a2, h2=h2)


def shift(p: np.ndarray, x: np.ndarray, lo: float = -0.24, hi: float = 0.66) -> np.ndarray:
    """(hi - lo) (sigmoid(z) - 1/2) in float64."""
    z, _ = forward(p, x)
    return (hi - lo) * (1.0 / (1.0 + np.exp(-z)) - 0.5)


def loss(p: np.ndarray, x: np.ndarray, y: np.ndarray, w: np.ndarray) -> float:
    z, _ = forward(p, x)
    sp = np.maximum(z, 0.0) + np.log1p(np.exp(-np.abs(z)))
    return float(np.mean(np.asarray(w, np.float64) * (sp - np.asarray(y, np.float64) * z)))


def grad(p: np.ndarray, x: np.ndarray, y: np.ndarray, w: np.ndarray) -> Tuple[np.ndarray, float]:
    """Gradient of loss() by the chain rule, and the loss."""
    q = unpack(p)
    z, c = forward(p, x)
    y = np.asarray(y, np.float64)
    w = np.asarray(w, np.float64)
    n = z.shape[0]
    sig = 1.0 / (1.0 + np.exp(-z))
    dz = w * (sig - y) / n
    g_w3 = c["h2"].T @ dz
    g_b3 = np.array([dz.sum()])
    d2 = np.outer(dz, q["w3"]) * (c["a2"] > 0)
    g_W2 = d2.T @ c["h1"]
    g_b2 = d2.sum(0)
    d1 = (d2 @ q["W2"]) * (c["a1"] > 0)
    g_W1 = d1.T @ c["x"]
    g_b1 = d1.sum(0)
    g = np.concatenate([g_W1.ravel(), g_b1, g_W2.ravel(), g_b2, g_w3, g_b3])
    return g, loss(p, x, y, w)


def bf16(a: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest, ties to even) and back, as the decode kernel
    rounds the network's operands for the matrix cores."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def split_bf16(a: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """The bf16 nearest each value and the bf16 nearest its remainder (the
    remainder of a float32 and its bf16 rounding is exact in float32)."""
    a = np.asarray(a, np.float32)
    hi = bf16(a)
    return hi, bf16(a - hi)


def shift_q8(p: np.ndarray, x: np.ndarray, lo: float = -0.24, hi: float = 0.66) -> np.ndarray:
    """Per-variable shift of a leg's draw as the decode kernel computes it at
    the start of relay leg 0, where feature 7 (the leg index) is 0: features 0-6
    in bf16 against W1 and b1 carried as bf16 pairs (nearest and remainder),
    first-layer activations in bf16, W2 in bf16 with b2 in fp32, second-layer
    activations and w3 in bf16, exact accumulation (rounded to fp32 after each
    layer), then 254 (sigmoid(z) - 1/2) in fp32 rounded to the nearest integer
    (ties to even) and scaled by (hi - lo) / 254.  The strength of the variable
    is clamp(u + shift, lo, hi) with u the leg's draw."""
    q = unpack(np.asarray(p, np.float32))
    xb = bf16(x[:, :7]).astype(np.float64)
    w1h, w1l = split_bf16(q["W1"][:, :7])
    b1h, b1l = split_bf16(q["b1"])
    w1 = w1h.astype(np.float64) + w1l.astype(np.float64)
    b1 = b1h.astype(np.float64) + b1l.astype(np.float64)
    h1 = bf16(np.maximum(xb @ w1.T + b1, 0.0).astype(np.float32)).astype(np.float64)
    h2 = bf16(np.maximum(h1 @ bf16(q["W2"]).astype(np.float64).T + q["b2"], 0.0).astype(np.float32))
    z = (h2.astype(np.float64) @ bf16(q["w3"]).astype(np.float64) + q["b3"][0]).astype(np.float32)
    sig = np.float32(1.0) / (np.float32(1.0) + np.exp(-z))
    steps = np.rint(np.float32(254.0) * (sig - np.float32(0.5))).astype(np.float32)
    scale = (np.float32(hi) - np.float32(lo)) * np.float32(1.0 / 254.0)
    return (steps * scale).astype(np.float32)
