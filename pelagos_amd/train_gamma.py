"""Train the learned memory strengths on the 936 x 8785 circuit-shaped code and
compare the learned form with uniform-draw legs.

    python -m pelagos_amd.train_gamma [--out PATH] [--train-shots 512] [--legs 4]
        [--samples 2000000] [--epochs 6] [--eval-shots 65536] [--learned-legs 4]
        [--method rwr|mean-draw]

Training shots have errors drawn at the workload's priors scaled to a mean of
0.003 (seed 1).  The default method (rwr, pelagos_amd.learned.rwr_training_set)
decodes them once per candidate shift with a constant network on the learned
legs and regresses the network onto the candidate shifts, each leg weighted by
how quickly it converged; features are those at the start of each learned leg
that starts from a valid solution (the only legs a network shifts).  The
evaluation decodes fresh shots (seed 2) drawn at the workload's own priors with
both forms and prints, per form, decoded shots/s (HIP-event kernel time and
wall time per 4096-shot batch), mean iterations, and the frame-error and
synthetic-observable error rates with Wilson 95% intervals.  One JSON line per
result.
"""

from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np
import torch

from . import RelayBPDecoder, codes, learned
from .observables import wilson

RELAY = dict(pre_iter=80, num_sets=100, set_max_iter=60, gamma0=0.65, gamma_dist_interval=(-0.24, 0.66),
             stop_nconv=5)


def evaluate(dec: RelayBPDecoder, H, err: np.ndarray, syn: np.ndarray, batch: int = 4096) -> dict:
    """Decode all shots in batches; time each batch's decode kernel (HIP events)
    and wall clock; count frame errors and synthetic-observable failures."""
    O = codes.synthetic_observables(H.shape[1], seed=0).astype(np.int64)
    dev = dec.device
    kern_ms, wall_s, iters, fe, le = 0.0, 0.0, 0, 0, 0
    dec.decode(torch.from_numpy(syn[:batch]).to(dev))  # warm-up (plans, gamma tables)
    torch.cuda.synchronize()
    for i in range(0, syn.shape[0], batch):
        s = torch.from_numpy(syn[i:i + batch]).to(dev)
        ev = []
        dec.kernel_events = ev
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = dec.decode(s)
        torch.cuda.synchronize()
        wall_s += time.perf_counter() - t0
        dec.kernel_events = None
        kern_ms += sum(a.elapsed_time(b) for a, b in ev)
        e = out["errors"].cpu().numpy()
        conv = out["found_count"].cpu().numpy() >= dec.stop_nconv
        iters += int(out["iterations"].cpu().numpy().astype(np.int64).sum())
        truth = err[i:i + batch]
        fe += int(((~conv) | (e != truth).any(1)).sum())
        resid = (e ^ truth).astype(np.int64)
        le += int(((~conv) | ((O @ resid.T) % 2).any(0)).sum())
    n = syn.shape[0]
    return dict(shots=n, shots_per_s_wall=round(n / wall_s, 1), shots_per_s_kernel=round(n / (kern_ms * 1e-3), 1),
                mean_iterations=round(iters / n, 3), frame_error=[round(x, 6) for x in wilson(fe, n)],
                synthetic_observable_ler=[round(x, 6) for x in wilson(le, n)], frame_errors=fe, logical_errors=le)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(learned.default_weights_path()))
    ap.add_argument("--train-shots", type=int, default=512)
    ap.add_argument("--legs", type=int, default=4)
    ap.add_argument("--samples", type=int, default=2_000_000)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval-shots", type=int, default=65536)
    ap.add_argument("--learned-legs", type=int, nargs="+", default=[4])
    ap.add_argument("--skip-train", action="store_true", help="evaluate the weights already at --out")
    ap.add_argument("--hard-weight", type=float, default=4.0,
                    help="weight of legs that started from a state violating some check")
    ap.add_argument("--easy-weight", type=float, default=1.0,
                    help="weight of legs that started from a valid solution (0 drops them)")
    ap.add_argument("--prior-mean", type=float, default=0.003,
                    help="mean prior of the training shots' error draws")
    ap.add_argument("--fast-iters", type=int, default=None,
                    help="a leg that starts from a valid solution counts as successful only when it "
                         "converges within this many iterations (default: any)")
    ap.add_argument("--save-legs", type=int, default=4, help="learned legs recorded with the weights")
    ap.add_argument("--method", choices=("rwr", "mean-draw"), default="rwr",
                    help="rwr: reward-weighted regression over constant candidate shifts; mean-draw: the "
                         "mean uniform draw of successful plain legs")
    ap.add_argument("--sigmas", type=float, nargs="+", default=[0.5, 2 / 3, 5 / 6, 0.99],
                    help="candidate network outputs (shift (hi - lo) (sigma - 1/2)) for rwr")
    ap.add_argument("--tau", type=float, default=1.0, help="rwr weight exp(-(leg iterations - 1) / tau)")
    ap.add_argument("--const-gamma", type=float, nargs="*", default=[],
                    help="also evaluate legs with this constant strength for every variable")
    args = ap.parse_args()
    log = lambda msg: print(msg, file=sys.stderr, flush=True)  # noqa: E731

    H, p = codes.banded_circuit_code(seed=0)
    out_path = Path(args.out)
    if not args.skip_train:
        scale = args.prior_mean / float(p.mean())
        err = codes.sample_errors(p, args.train_shots, seed=1, scale=scale)
        syn = codes.syndromes_of(H, err)
        t0 = time.perf_counter()
        if args.method == "rwr":
            def make_decoder(params, legs):
                return RelayBPDecoder(H, p, data_type="float32", algo="relay", seed=1, device=[0],
                                      learned_gamma=params, learned_legs=legs, **RELAY)

            ts = learned.rwr_training_set(make_decoder, syn, err, args.legs, sigmas=args.sigmas, tau=args.tau,
                                          max_samples=args.samples, seed=1)
            info = {k: ts[k] for k in ("shots_shifted", "shots_correct", "mean_leg_iterations", "candidates")}
            info.update(method="rwr", sigmas=args.sigmas, tau=args.tau)
        else:
            dec = RelayBPDecoder(H, p, data_type="float32", algo="relay", seed=1, device=[0], **RELAY)
            ts = learned.training_set(dec, syn, args.legs, args.samples, seed=1, hard_weight=args.hard_weight,
                                      easy_weight=args.easy_weight, fast_iters=args.fast_iters)
            info = {k: ts[k] for k in ("legs_started", "legs_converged", "converged_from_violated", "candidates")}
            info.update(method="mean-draw", hard_weight=args.hard_weight, easy_weight=args.easy_weight,
                        fast_iters=args.fast_iters)
        log(f"[train] {ts['x'].shape[0]} samples in {time.perf_counter() - t0:.1f} s: {info}")
        net = learned.GammaMLP(seed=1)
        t0 = time.perf_counter()
        losses = net.fit(ts["x"], ts["y"], ts["w"], epochs=args.epochs, lr=args.lr, log=log)
        torch.cuda.synchronize()
        train_s = time.perf_counter() - t0
        out_path.parent.mkdir(parents=True, exist_ok=True)
        net.save(out_path, learned_legs=args.save_legs)
        print(json.dumps({"train": {"samples": int(ts["x"].shape[0]), **info, "epochs": args.epochs,
                                    "losses": [round(x, 5) for x in losses], "seconds": round(train_s, 2),
                                    "prior_scale": round(scale, 4), "weights": str(out_path)}}), flush=True)
    net = learned.GammaMLP.load(out_path)

    err = codes.sample_errors(p, args.eval_shots, seed=2)
    syn = codes.syndromes_of(H, err)
    plain = RelayBPDecoder(H, p, data_type="float32", algo="relay", seed=2, device=[0], **RELAY)
    res = {"plain": evaluate(plain, H, err, syn)}
    print(json.dumps({"eval": "plain", **res["plain"]}), flush=True)
    for cg in args.const_gamma:
        eg = np.full((1, H.shape[1]), cg, np.float32)
        dec = RelayBPDecoder(H, p, data_type="float32", algo="relay", seed=2, device=[0], explicit_gammas=eg,
                             **RELAY)
        r = evaluate(dec, H, err, syn)
        r["speedup_kernel"] = round(r["shots_per_s_kernel"] / res["plain"]["shots_per_s_kernel"], 3)
        print(json.dumps({"eval": f"constant_gamma={cg}", **r}), flush=True)
    for L in args.learned_legs:
        dec = RelayBPDecoder(H, p, data_type="float32", algo="relay", seed=2, device=[0], learned_gamma=net,
                             learned_legs=L, **RELAY)
        r = evaluate(dec, H, err, syn)
        r["speedup_kernel"] = round(r["shots_per_s_kernel"] / res["plain"]["shots_per_s_kernel"], 3)
        lo, hi = res["plain"]["synthetic_observable_ler"][1:]
        r["ler_within_plain_ci"] = bool(lo <= r["synthetic_observable_ler"][0] <= hi)
        print(json.dumps({"eval": f"learned_legs={L}", **r}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
