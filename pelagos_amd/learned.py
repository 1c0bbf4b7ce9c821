"""Learned per-variable memory strengths for the first relay legs.

A small network (include/pelagos_relaybp.h: 8 -> 32 -> 32 -> 1) proposes the
memory strength of every variable at the start of each of the first
``learned_legs`` relay legs, from features of the state the previous leg ended
in.  The network does not replace a leg's random draw, whose disorder is what
lets successive legs explore differently (a fixed strength per variable
repeats a failed leg forever); it shifts each variable's draw by
(hi - lo) (sigmoid(z) - 1/2).  It is trained by reward-weighted regression
(``rwr_training_set``): the training shots are decoded once per candidate
constant shift, and each shot whose pre-iter leg ended at a valid solution
contributes its variables' features with the candidate as target, weighted by
how short its learned legs were when it still decoded to the true error.  The
fitted network so proposes, per state, the shift under which the relay legs
converge fastest without changing the decoding.  (``training_set``, the
earlier mean-draw method, regresses onto the mean uniform draw of legs that
converged; it is kept for comparison.)  Training runs on the GPU through the
library's own forward/backward and Adam kernels (``pel_mlp_grad``,
``pel_mlp_adam``); at decode time the network runs inside the decode kernel on
the matrix cores.

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

    def shift(self, x: torch.Tensor, lo: float = -0.24, hi: float = 0.66) -> torch.Tensor:
        """The shift (hi - lo) (sigmoid(z) - 1/2) a learned leg adds to each
        variable's uniform draw (in fp32; the decode kernel's inference rounds
        operands to bf16 and the shift to a multiple of (hi - lo) / 254)."""
        return (hi - lo) * (torch.sigmoid(self.logits(x)) - 0.5)

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
    def save(self, path, learned_legs: Optional[int] = None, learned_last_damp: Optional[float] = None) -> None:
        extra = {} if learned_legs is None else {"learned_legs": np.int32(learned_legs)}
        if learned_last_damp is not None:
            extra["learned_last_damp"] = np.float32(learned_last_damp)
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
    saved, saved_l = dec.num_sets, dec.learned_legs
    tot = []
    try:
        for r in range(legs + 1):
            dec.num_sets = min(r, saved)
            dec.learned_legs = min(saved_l, dec.num_sets)   # learned legs among the first r
            tot.append(dec.decode(syndromes)["iterations"].cpu().numpy().astype(np.int64))
    finally:
        dec.num_sets, dec.learned_legs = saved, saved_l
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


def constant_params(sigma: float) -> np.ndarray:
    """A network whose output is sigmoid(z) = sigma for every input (every
    weight 0, b3 = logit(sigma)): it shifts every variable's draw by
    (hi - lo) (sigma - 1/2)."""
    p = np.zeros(N_PARAMS, np.float32)
    p[-1] = np.float32(np.log(sigma / (1.0 - sigma)))
    return p


def rwr_training_set(make_decoder, syndromes: np.ndarray, errors: np.ndarray, legs: int,
                     sigmas=(0.5, 2 / 3, 5 / 6, 0.99), tau: float = 1.0, max_samples: int = 2_000_000,
                     seed: int = 1, batch: int = 512) -> Dict[str, torch.Tensor]:
    """Reward-weighted regression over candidate shifts.  For each candidate
    sigma, ``make_decoder(params, legs)`` gives a decoder whose network is the
    constant ``sigma``: every shot whose pre-iter leg ended at a valid solution
    has the draws of its learned relay legs 0..legs-1 shifted by
    (hi - lo) (sigma - 1/2), as in every learned decode (the network runs once,
    on the state after the pre-iter leg).  Such a shot contributes the features
    of its variables at that state, with target sigma and weight
    exp(-(m - 1) / tau), m the mean length of the learned legs it ran
    (``leg_iterations``), times 1 if it was decoded to ``errors`` (the true
    error pattern) and 0 otherwise.  The fitted sigmoid is then, for each
    state, the reward-weighted mean of the candidates: the shift under which
    legs from such states converge fastest without changing the decoding.  A
    uniform sample of at most ``max_samples`` variables is kept."""
    xs, ys, ws = [], [], []
    stats = {"shots_shifted": 0, "shots_correct": {}, "mean_leg_iterations": {}}
    for sg in sigmas:
        dec = make_decoder(constant_params(sg), legs)
        dev = dec.device
        H = dec.H_csr.tocsr().astype(np.int32)
        n_it, n_leg, n_ok, n_sh = 0, 0, 0, 0
        for i in range(0, syndromes.shape[0], batch):
            syn = syndromes[i:i + batch]
            syn_dev = torch.from_numpy(np.ascontiguousarray(syn)).to(dev)
            out = dec.collect_traces(syn_dev, legs, with_network=True)
            hard = out["hard"].cpu().numpy()[:, 0]                         # [b, V] at relay leg 0
            reached = hard[:, 0] != 255
            shifted = np.zeros_like(reached)
            shifted[reached] = _syndrome_ok(H, np.where(hard[:, None, :] == 255, 0, hard[:, None, :]),
                                            syn)[reached, 0]
            correct = (out["errors"].cpu().numpy() == errors[i:i + batch]).all(1)
            its = leg_iterations(dec, syn_dev, legs)                       # [b, legs]
            ran = np.maximum((its > 0).sum(1), 1)
            mean_it = its.sum(1) / ran
            w = np.where(shifted & correct, np.exp(-(mean_it - 1.0) / tau), 0.0).astype(np.float32)
            idx = np.nonzero(shifted)[0]
            n_sh += int(idx.size)
            n_ok += int((shifted & correct).sum())
            n_it += int(its[idx].sum())
            n_leg += int(ran[idx].sum())
            if idx.size == 0:
                continue
            f = out["features"][torch.from_numpy(idx).to(dev), 0]          # [k, V, 8]
            wt = torch.from_numpy(w[idx]).to(dev)
            xs.append(f.reshape(-1, N_FEATURES))
            ys.append(torch.full((f.shape[0] * f.shape[1],), float(sg), dtype=torch.float32, device=dev))
            ws.append(wt[:, None].expand(-1, f.shape[1]).reshape(-1))
        key = round(float(sg), 4)
        stats["mean_leg_iterations"][key] = round(n_it / max(n_leg, 1), 3)
        stats["shots_correct"][key] = n_ok
        stats["shots_shifted"] = n_sh
        del dec
    x, y, w = torch.cat(xs), torch.cat(ys), torch.cat(ws)
    keep_mask = w > 0
    x, y, w = x[keep_mask], y[keep_mask], w[keep_mask]
    gen = torch.Generator(device="cpu").manual_seed(seed)
    keep = torch.randperm(x.shape[0], generator=gen)[:max_samples].to(x.device)
    w = w[keep]
    w = w / w.mean()
    stats["candidates"] = int(keep_mask.sum())
    return dict(x=x[keep].contiguous(), y=y[keep].contiguous(), w=w.contiguous(), **stats)


def default_weights_path() -> Path:
    return WEIGHTS_DIR / "gamma_mlp_banded936x8785.npz"
