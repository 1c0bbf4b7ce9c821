"""Learned per-variable memory strengths for the first relay legs.

A small network (include/pelagos_relaybp.h: 8 -> 32 -> 32 -> 1) proposes the
memory strength of every variable at the start of each of the first
``learned_legs`` relay legs, from features of the state the previous leg ended
in.  It is trained on traces of the plain decoder (uniform-draw legs): the
target is whether a variable's hard decision is already correct, so
confident-correct variables keep strong memory and suspect ones get
disordered memory.  Training runs on the GPU through the library's own
forward/backward and Adam kernels (``pel_mlp_grad``, ``pel_mlp_adam``); at
decode time the network runs inside the decode kernel on the matrix cores.

Uniform-draw legs remain the fallback (every leg after the learned ones), and
the plain decoder's acceptance rules judge every shot: a solution counts only
if it satisfies every check, and the lowest weight wins.

The reference has no learned component; its plug points for strengths are
``explicit_gammas`` and ``init_gamma`` (decoder.py:480-491, 2322).
"""

from __future__ import annotations

import ctypes
from pathlib import Path
from typing import Dict, Optional

import numpy as np
import torch

from . import _native

N_FEATURES = 8
HIDDEN = 32
N_PARAMS = _native.PEL_MLP_PARAMS
GAMMA_LO, GAMMA_SPAN = -0.24, 0.9
FEATURE_NAMES = ("prior_llr/8", "posterior/16", "flips/8", "mean_incoming/8", "min_incoming/8",
                 "column_weight/8", "violated_check_fraction", "leg/4")
WEIGHTS_DIR = Path(__file__).resolve().parent / "data"


def init_params(seed: int = 0) -> np.ndarray:
    """He-initialised weights, zero biases, and b3 = log 9 so that every
    strength starts near sigmoid = 0.9 ('keep memory')."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0.0, np.sqrt(2.0 / N_FEATURES), (HIDDEN, N_FEATURES))
    W2 = rng.normal(0.0, np.sqrt(2.0 / HIDDEN), (HIDDEN, HIDDEN))
    w3 = rng.normal(0.0, np.sqrt(1.0 / HIDDEN), HIDDEN)
    p = np.concatenate([W1.ravel(), np.zeros(HIDDEN), W2.ravel(), np.zeros(HIDDEN), w3, [np.log(9.0)]])
    return p.astype(np.float32)


class GammaMLP:
    """The network's parameters on one device, with the library's training
    kernels."""

    def __init__(self, params=None, device="cuda:0", seed: int = 0):
        self.device = torch.device(device)
        p = init_params(seed) if params is None else np.asarray(params, np.float32).reshape(-1)
        if p.size != N_PARAMS:
            raise ValueError(f"expected {N_PARAMS} parameters, got {p.size}")
        self.params = torch.from_numpy(p.copy()).to(self.device)
        self._m1 = torch.zeros_like(self.params)
        self._m2 = torch.zeros_like(self.params)
        self._step = 0
        self._ws: Optional[torch.Tensor] = None

    # -- kernels -------------------------------------------------------------------------------
    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def logits(self, x: torch.Tensor) -> torch.Tensor:
        x = x.to(self.device, torch.float32).contiguous()
        z = torch.empty(x.shape[0], dtype=torch.float32, device=self.device)
        _native.check(_native.lib().pel_mlp_forward(self.params.data_ptr(), x.data_ptr(), int(x.shape[0]),
                                                    z.data_ptr(), self._stream()), "pel_mlp_forward", mlp=True)
        return z

    def gamma(self, x: torch.Tensor) -> torch.Tensor:
        return GAMMA_LO + GAMMA_SPAN * torch.sigmoid(self.logits(x))

    def grad(self, x: torch.Tensor, y: torch.Tensor, w: torch.Tensor):
        """(gradient of the mean weighted logistic loss, loss) from pel_mlp_grad."""
        L = _native.lib()
        nbytes = int(L.pel_mlp_grad_workspace_bytes())
        if self._ws is None or self._ws.numel() < nbytes:
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        g = torch.empty_like(self.params)
        loss = torch.empty(1, dtype=torch.float32, device=self.device)
        x = x.to(self.device, torch.float32).contiguous()
        y = y.to(self.device, torch.float32).contiguous()
        w = w.to(self.device, torch.float32).contiguous()
        _native.check(L.pel_mlp_grad(self.params.data_ptr(), x.data_ptr(), y.data_ptr(), w.data_ptr(),
                                     int(x.shape[0]), g.data_ptr(), loss.data_ptr(), self._ws.data_ptr(), nbytes,
                                     self._stream()), "pel_mlp_grad", mlp=True)
        return g, loss

    def adam_step(self, g: torch.Tensor, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999,
                  eps: float = 1e-8) -> None:
        self._step += 1
        _native.check(_native.lib().pel_mlp_adam(self.params.data_ptr(), self._m1.data_ptr(), self._m2.data_ptr(),
                                                 g.data_ptr(), float(lr), float(beta1), float(beta2), float(eps),
                                                 self._step, self._stream()), "pel_mlp_adam", mlp=True)

    def fit(self, x: torch.Tensor, y: torch.Tensor, w: torch.Tensor, epochs: int = 4, batch: int = 16384,
            lr: float = 2e-3, seed: int = 0, log=None) -> list:
        """Mini-batch Adam over shuffled samples; returns the mean loss per epoch."""
        n = x.shape[0]
        gen = torch.Generator(device="cpu").manual_seed(seed)
        losses = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=gen).to(self.device)
            tot, cnt = torch.zeros((), device=self.device), 0
            for i in range(0, n, batch):
                idx = perm[i:i + batch]
                g, loss = self.grad(x[idx], y[idx], w[idx])
                self.adam_step(g, lr=lr)
                tot += loss[0] * idx.numel()
                cnt += idx.numel()
            losses.append(float(tot) / cnt)
            if log:
                log(f"epoch {ep}: loss {losses[-1]:.5f}")
        return losses

    # -- persistence ---------------------------------------------------------------------------
    def save(self, path) -> None:
        np.savez(path, params=self.params.cpu().numpy())

    @classmethod
    def load(cls, path, device="cuda:0") -> "GammaMLP":
        with np.load(path, allow_pickle=False) as f:
            return cls(f["params"], device=device)


def training_set(dec, errors: np.ndarray, syndromes: np.ndarray, legs: int, max_samples: int, seed: int,
                 batch: int = 256) -> Dict[str, torch.Tensor]:
    """Traces of the plain decoder ``dec`` on the given shots: features at the
    start of relay legs 0..legs-1 with target 1 where the hard decision equals
    the true error bit.  Keeps every incorrect variable and a uniform sample
    of the correct ones, up to ``max_samples`` in total; the correct ones are
    re-weighted to stand for all of them."""
    dev = dec.device
    xs, ys = [], []
    n_correct_total = 0
    for i in range(0, syndromes.shape[0], batch):
        syn = torch.from_numpy(syndromes[i:i + batch]).to(dev)
        out = dec.collect_traces(syn, legs)
        hard = out["hard"]                                     # [b, legs, V]
        truth = torch.from_numpy(errors[i:i + batch]).to(dev)[:, None, :].expand_as(hard)
        ok = hard != 255
        correct = (hard == truth) & ok
        feats = out["features"][ok]
        ys.append(correct[ok].to(torch.float32))
        xs.append(feats)
        n_correct_total += int(correct.sum())
    x = torch.cat(xs)
    y = torch.cat(ys)
    wrong = torch.nonzero(y == 0).flatten()
    right = torch.nonzero(y == 1).flatten()
    keep_right = max(0, min(right.numel(), max_samples - wrong.numel()))
    gen = torch.Generator(device="cpu").manual_seed(seed)
    pick = right[torch.randperm(right.numel(), generator=gen)[:keep_right].to(dev)]
    idx = torch.cat([wrong, pick])
    w = torch.ones(idx.numel(), dtype=torch.float32, device=dev)
    # correct samples stand for all correct variables; incorrect ones are up-weighted
    # so that the two classes carry comparable total weight (capped)
    scale_right = right.numel() / max(keep_right, 1)
    w[wrong.numel():] = 1.0
    ratio = min(50.0, (right.numel() / max(wrong.numel(), 1)) / scale_right)
    w[:wrong.numel()] = ratio
    return dict(x=x[idx].contiguous(), y=y[idx].contiguous(), w=w, n_wrong=int(wrong.numel()),
                n_right=int(right.numel()))


def default_weights_path() -> Path:
    return WEIGHTS_DIR / "gamma_mlp_banded936x8785.npz"
