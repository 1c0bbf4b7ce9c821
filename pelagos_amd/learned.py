"""Learned per-variable memory strengths for the first relay legs.

A small network (include/pelagos_relaybp.h: 8 -> 32 -> 32 -> 1) proposes the
memory strength of every variable at the start of each of the first
``learned_legs`` relay legs, from features of the state the previous leg ended
in.  The network does not replace a leg's random draw, whose disorder is what
lets successive legs explore differently (a fixed strength per variable
repeats a failed leg forever); it shifts each variable's draw by
(hi - lo) (sigmoid(z) - 1/2).  It is trained on traces of the plain decoder
(uniform-draw legs) to predict, from the features, the mean draw of the legs
that converged, so it moves each variable's strengths towards the values that
let legs converge.  Training runs on the GPU through the library's own
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
    def save(self, path, learned_legs: Optional[int] = None) -> None:
        extra = {} if learned_legs is None else {"learned_legs": np.int32(learned_legs)}
        np.savez(path, params=self.params.cpu().numpy(), **extra)

    @classmethod
    def load(cls, path, device="cuda:0") -> "GammaMLP":
        with np.load(path, allow_pickle=False) as f:
            return cls(f["params"], device=device)


def _syndrome_ok(H, hard: np.ndarray, syn: np.ndarray) -> np.ndarray:
    """[n, k] bool: row hard[i, j] satisfies H h = syn[i]."""
    n, k, V = hard.shape
    s = (H @ hard.reshape(n * k, V).T.astype(np.int32)) % 2            # [C, n k]
    return (s.T.reshape(n, k, -1) == syn[:, None, :]).all(2)


def leg_iterations(dec, syndromes: torch.Tensor, legs: int) -> np.ndarray:
    """[b, legs] BP iterations of relay legs 0 .. legs-1 of every shot under
    ``dec`` (0 for a leg the shot never started): the differences of the
    iteration counts of decodes stopped after r and r + 1 relay legs, whose
    legs are the full decode's own (each leg's draw depends on the leg only)."""
    saved = dec.num_sets
    tot = []
    try:
        for r in range(legs + 1):
            dec.num_sets = min(r, saved)
            tot.append(dec.decode(syndromes)["iterations"].cpu().numpy().astype(np.int64))
    finally:
        dec.num_sets = saved
    return np.diff(np.stack(tot, 1), axis=1)


def training_set(dec, syndromes: np.ndarray, legs: int, max_samples: int, seed: int, batch: int = 256,
                 hard_weight: float = 4.0, easy_weight: float = 1.0,
                 fast_iters: Optional[int] = None) -> Dict[str, torch.Tensor]:
    """Traces of the plain decoder ``dec`` (uniform-draw legs) on the given
    shots.  For every relay leg r < legs that a shot started: the features of
    every variable at the leg's start, and as target the variable's draw u
    of that leg mapped to [0, 1], (u - lo) / (hi - lo).  Only successful legs
    count, so the fitted sigmoid is the mean draw of successful legs given
    the features.  A leg that started from a state violating some check
    (weight ``hard_weight``) succeeds when it converged (the next leg starts
    from a valid solution, or the shot stopped after it); a leg that started
    from a valid solution (weight ``easy_weight``, 0 drops them) succeeds when
    it converged within ``fast_iters`` iterations (any number when None).  A
    uniform sample of at most ``max_samples`` variables is kept."""
    dev = dec.device
    H = dec.H_csr.tocsr().astype(np.int32)
    lo, hi = (float(x) for x in dec.gamma_dist_interval)
    rows = dec.gamma_table(1).cpu().numpy()                              # [R, V] caller order
    xs, ys, ws = [], [], []
    n_legs = n_ok = n_hard = 0
    for i in range(0, syndromes.shape[0], batch):
        syn = syndromes[i:i + batch]
        syn_dev = torch.from_numpy(syn).to(dev)
        out = dec.collect_traces(syn_dev, legs + 1)
        hard = out["hard"].cpu().numpy()                                 # [b, legs+1, V]
        reached = hard[:, :, 0] != 255
        valid = np.zeros_like(reached)
        valid[reached] = _syndrome_ok(H, np.where(hard == 255, 0, hard), syn)[reached]
        its = leg_iterations(dec, syn_dev, legs) if fast_iters is not None else None
        for r in range(legs):
            started = reached[:, r]
            ok = started & (valid[:, r + 1] | ~reached[:, r + 1])
            from_bad = started & ~valid[:, r]
            if its is not None:
                ok &= from_bad | (its[:, r] <= fast_iters)
            if easy_weight <= 0:
                ok &= from_bad
            n_legs += int(started.sum())
            n_ok += int(ok.sum())
            n_hard += int((ok & from_bad).sum())
            idx = np.nonzero(ok)[0]
            if idx.size == 0:
                continue
            f = out["features"][torch.from_numpy(idx).to(dev), r]          # [k, V, 8]
            y = torch.from_numpy(((rows[r] - lo) / (hi - lo)).astype(np.float32)).to(dev)
            w = torch.from_numpy(np.where(from_bad[idx], hard_weight, easy_weight).astype(np.float32)).to(dev)
            xs.append(f.reshape(-1, N_FEATURES))
            ys.append(y[None, :].expand(idx.size, -1).reshape(-1))
            ws.append(w[:, None].expand(-1, dec.V).reshape(-1))
    x, y, w = torch.cat(xs), torch.cat(ys), torch.cat(ws)
    gen = torch.Generator(device="cpu").manual_seed(seed)
    keep = torch.randperm(x.shape[0], generator=gen)[:max_samples].to(dev)
    return dict(x=x[keep].contiguous(), y=y[keep].contiguous(), w=w[keep].contiguous(), legs_started=n_legs,
                legs_converged=n_ok, converged_from_violated=n_hard, candidates=int(x.shape[0]))


def default_weights_path() -> Path:
    return WEIGHTS_DIR / "gamma_mlp_banded936x8785.npz"
