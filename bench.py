"""Decoded shots/s of Relay-BP-S on one MI355X (936 x 8785 banded circuit-shaped code),
plain and with learned strengths, plus serving latency.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--shots 4096] [--depth 3] [--requeue-below 0]
                    [--no-cpu] [--no-learned] [--no-damped] [--no-serving] [--no-batched] [--no-bb72]

A step decodes the whole 4096-shot batch (syndromes already resident in HBM)
into its own output buffers.  The timed steps run with ``--depth`` launches in
flight: step i is issued on stream i % depth (each stream with its own
workspace and outputs), so a launch's last few slow shots overlap the next
launches instead of leaving the GPU idle.  With launches overlapping, the
timed steps skip the library's re-queue of fast first legs (a single launch's
tail remedy, which parks and reloads every such shot): ``--requeue-below``
sets the plan's threshold for them (default 0, off).  The one-launch-at-a-time
figures use the default plan.  Every step's results are checked
against a plain ``decode()`` of the batch.  Prints ONE JSON line (rank 0)
with: the throughput of the plain form (uniform-draw legs) and, when trained
weights are present, of the learned form on the same shots with both error
rates; ``value`` is the faster form whose synthetic-observable error rate
lies in the plain form's 95% interval.  A learned variant with a fourth
learned leg at half shift (not the default) is reported beside them.  Also the decode kernel's roofline
figures (algorithmic edge-message bytes over the per-launch kernel time
measured with HIP events on each launch's stream, the aggregate rate, and the
same for launches one at a time), the [[72,12,6]] code at code capacity
(shots/s, LDS rate, per-shot latency, logical error rate), one graph launch
over three codes, the
continuous-batching server's p50/p99 latency at five Poisson arrival rates,
and the CPU oracle timed on the same shots on this host.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
WORKLOAD = dict(n_blocks=13, checks_per_block=72, V=8785, seed=0)
RELAY = dict(pre_iter=80, num_sets=100, set_max_iter=60, gamma0=0.65,
             gamma_dist_interval=(-0.24, 0.66), stop_nconv=5)


def wilson(k: int, n: int, z: float = 1.959964):
    if n == 0:
        return (0.0, 0.0, 1.0)
    p = k / n
    den = 1 + z * z / n
    c = (p + z * z / (2 * n)) / den
    h = z * math.sqrt(p * (1 - p) / n + z * z / (4 * n * n)) / den
    return (p, max(0.0, c - h), min(1.0, c + h))


def bytes_per_iteration(C: int, V: int, E: int) -> int:
    """Algorithmic HBM bytes of one BP iteration of one shot with edge messages in
    HBM (fp32): read mu + write nu (C2V), read nu + write mu (V2C), read+write the
    posterior, bit-packed hard decisions and syndrome."""
    return 16 * E + 8 * V + (V + 7) // 8 + (C + 7) // 8


def cpu_baseline(dec, syn_np, lam, gtab, target_s: float = 10.0):
    """The C oracle (one shot per OpenMP thread, all host cores) on the same
    shots and gamma table, repeated over the batch until about target_s of CPU
    work; plus a single-thread figure on a smaller prefix."""
    import numpy as np

    import oracle

    threads = os.cpu_count() or 1
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    kw = dict(pre_iter=RELAY["pre_iter"], num_sets=RELAY["num_sets"],
              set_max_iter=RELAY["set_max_iter"], stop_nconv=RELAY["stop_nconv"],
              gamma0=RELAY["gamma0"])
    g = dec.graph
    B = syn_np.shape[0]
    done, reps, ref = 0, 0, None
    t0 = time.perf_counter()
    while True:
        out = oracle.decode(g.chk_ptr, g.edge_var, g.V, syn_np, lam, gtab, nthreads=threads, **kw)
        ref = ref or out
        done += B
        reps += 1
        dt = time.perf_counter() - t0
        if dt >= target_s or reps >= 50:
            break
    n1 = min(B, 256)
    t1 = time.perf_counter()
    oracle.decode(g.chk_ptr, g.edge_var, g.V, syn_np[:n1], lam, gtab, nthreads=1, **kw)
    dt1 = time.perf_counter() - t1
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), "")
    except OSError:
        pass
    return {
        "value": round(done / dt, 1),
        "unit": "shots/s",
        "cores": threads,
        "nproc": os.cpu_count(),
        "cpu_model": model,
        "kind": "port",
        "sample": f"all {B} shots x {reps} passes ({dt:.1f} s), C oracle, OpenMP one shot per thread; "
                  f"single thread: {n1 / dt1:.1f} shots/s on the first {n1} shots",
        "single_thread_value": round(n1 / dt1, 1),
    }, ref, B


def _pmc_rows(folder: str):
    import csv
    import glob

    rows = []
    for f in sorted(glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True)):
        with open(f, newline="") as fh:
            rows += list(csv.DictReader(fh))
    return rows


def traffic_record(fetch_dir: str, write_dir: str, cfetch_dir: str, cwrite_dir: str) -> int:
    """HBM bytes per decode launch from FETCH_SIZE / WRITE_SIZE (KiB per
    dispatch, averaged over the run's launches of each kernel variant: plain
    and learned), each scaled by what the same counter reads for a device copy
    of known size (the copy reads and writes copy_probe_bytes)."""
    def kib(rows, pred, counter):
        return [float(r["Counter_Value"]) for r in rows if pred(r["Kernel_Name"]) and r["Counter_Name"] == counter]

    copy_bytes = 1024 << 20
    cp_f = kib(_pmc_rows(cfetch_dir), lambda k: "copyBuffer" in k, "FETCH_SIZE")
    cp_w = kib(_pmc_rows(cwrite_dir), lambda k: "copyBuffer" in k, "WRITE_SIZE")
    f_scale = copy_bytes / (sum(cp_f) / len(cp_f) * 1024)
    w_scale = copy_bytes / (sum(cp_w) / len(cp_w) * 1024)
    rec = {"kernel": "relay_decode_kernel", "workload": "banded 936x8785, 4096 shots, one launch",
           "copy_probe_bytes": copy_bytes, "copy_fetch_kib_raw": sum(cp_f) / len(cp_f),
           "copy_write_kib_raw": sum(cp_w) / len(cp_w), "fetch_scale": round(f_scale, 4),
           "write_scale": round(w_scale, 4)}
    frows, wrows = _pmc_rows(fetch_dir), _pmc_rows(write_dir)
    for form, flag in (("plain", "false>"), ("learned", "true>")):
        pred = (lambda k, flag=flag: "relay_decode_kernel" in k and flag in k)
        f, w = kib(frows, pred, "FETCH_SIZE"), kib(wrows, pred, "WRITE_SIZE")
        if not f or not w:
            continue
        rec[form] = {"launches": len(f), "fetch_kib_raw": sum(f) / len(f), "write_kib_raw": sum(w) / len(w),
                     "hbm_bytes_per_launch": int(sum(f) / len(f) * 1024 * f_scale + sum(w) / len(w) * 1024 * w_scale)}
    out = ROOT / "profiles" / "relay_decode_banded936x8785_traffic.json"
    out.write_text(json.dumps(rec, indent=1) + "\n")
    print(json.dumps(rec))
    return 0


def roofline(iters_per_launch: int, b_iter: int, kern_ms, elapsed: float, steps: int, seq_ms: float, depth: int,
             traffic, kernel: str) -> dict:
    """The decode kernel's roofline object: algorithmic bytes per launch
    (B_iter x shot-iterations) over the average launch duration (HIP events
    on each launch's stream) in the timed steps, beside the aggregate rate
    over the timed region and the rate of launches one at a time."""
    kern_avg = sum(kern_ms) / max(len(kern_ms), 1)
    bytes_launch = iters_per_launch * b_iter
    achieved = bytes_launch / (kern_avg * 1e-3) / 1e9
    aggregate = bytes_launch * steps / elapsed / 1e9
    isolated = bytes_launch / (seq_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kernel,
            "kernel_ms": round(kern_avg, 4), "launches_in_flight": depth, "bytes_per_iteration": b_iter,
            "iterations_per_launch": iters_per_launch, "aggregate_achieved": round(aggregate, 1),
            "aggregate_frac": round(aggregate / HBM_PEAK_GBS, 4), "isolated_kernel_ms": round(seq_ms, 4),
            "isolated_achieved": round(isolated, 1), "isolated_frac": round(isolated / HBM_PEAK_GBS, 4)}


def copy_bandwidth(mib: int = 1024, reps: int = 10) -> float:
    """Achievable HBM bandwidth on this GPU: a device-to-device copy of ``mib``
    MiB, read plus written bytes over its HIP-event time (GB/s)."""
    import torch

    n = mib << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    b.synchronize()
    gbs = 2.0 * n * reps / (a.elapsed_time(b) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return round(gbs, 1)


def _traffic(form: str):
    tfile = ROOT / "profiles" / "relay_decode_banded936x8785_traffic.json"
    try:
        rec = json.loads(tfile.read_text())
        return rec.get(form, {}).get("hbm_bytes_per_launch")
    except Exception:
        return None


def measure(dec, syn, steps: int, warmup: int, depth: int) -> dict:
    """Time ``steps`` decodes of ``syn`` with ``depth`` launches in flight
    (step i on stream i % depth, each stream with its own workspace and output
    buffers), after ``warmup`` untimed ones, bracketed by device
    synchronisation.  Returns the rate, the per-launch kernel times (HIP events
    recorded by the library around each decode kernel on its stream), and the
    outputs of the last step."""
    import torch

    B, V = int(syn.shape[0]), dec.V
    dev = syn.device
    wsb = dec.prepare(B)
    streams = [torch.cuda.Stream(device=dev) for _ in range(depth)]
    slots = [dict(ws=torch.empty(wsb, dtype=torch.uint8, device=dev),
                  errors=torch.empty(B, V, dtype=torch.uint8, device=dev),
                  weights=torch.empty(B, dtype=torch.float32, device=dev),
                  iterations=torch.empty(B, dtype=torch.int32, device=dev),
                  found_count=torch.empty(B, dtype=torch.int32, device=dev)) for _ in range(depth)]

    def issue(n, events):
        cur = torch.cuda.current_stream(dev)
        for st in streams:
            st.wait_stream(cur)
        dec.kernel_events = events
        for i in range(n):
            o = slots[i % depth]
            dec.decode_async(syn, o["errors"], o["weights"], o["iterations"], o["found_count"], o["ws"],
                             streams[i % depth])
        dec.kernel_events = None
        for st in streams:
            cur.wait_stream(st)

    issue(max(warmup, depth), None)
    torch.cuda.synchronize(dev)
    ev = []
    t0 = time.perf_counter()
    issue(steps, ev)
    torch.cuda.synchronize(dev)
    elapsed = time.perf_counter() - t0
    kern = [a.elapsed_time(b) for a, b in ev]
    return {"rate": B * steps / elapsed, "elapsed": elapsed, "kernel_ms": kern,
            "out": slots[(steps - 1) % depth], "slots": slots[:min(depth, steps)]}


def ler_eval(dec, H, p, batches: int = 16, shots: int = 4096) -> dict:
    """Synthetic-observable logical error rate of ``dec`` on ``batches`` fresh
    batches of ``shots`` shots (errors drawn from the priors, seeds 101..), a
    shot failing when it did not converge or its decoding differs from the
    true error by an observable; Wilson 95% interval."""
    import numpy as np
    import torch

    from pelagos_amd import codes

    O = codes.synthetic_observables(H.shape[1], seed=0).astype(np.int64)
    fails = n = 0
    for k in range(batches):
        err = codes.sample_errors(p, shots, seed=101 + k)
        syn = torch.from_numpy(codes.syndromes_of(H, err)).to(dec.device).contiguous()
        out = dec.decode(syn)
        e = out["errors"].cpu().numpy()
        conv = out["found_count"].cpu().numpy() >= dec.stop_nconv
        fails += int(((~conv) | ((O @ (e ^ err).T.astype(np.int64)) % 2).any(0)).sum())
        n += shots
    lo_hi = wilson(fails, n)
    return {"shots": n, "failures": fails, "ler": [round(x, 6) for x in lo_hi]}


def same_outputs(slots, ref) -> bool:
    import torch

    return all(torch.equal(o[k], ref[k]) for o in slots for k in ("errors", "weights", "iterations", "found_count"))


def learned_form(H, p, err, syn, syn_np, args, plain_rate, plain_ler, legs=None, damp=None) -> dict:
    """The learned-strength form on the same shots (same steps/warmup).  ``legs``
    and ``damp`` override the learned-leg count and the last learned leg's
    damping stored with the weights."""
    import numpy as np
    import torch

    from pelagos_amd import RelayBPDecoder, codes, learned

    wpath = learned.default_weights_path()
    if not wpath.exists():
        return {"skipped": f"no trained weights at {wpath.relative_to(ROOT)}"}
    with np.load(wpath, allow_pickle=False) as f:
        stored_legs = int(f["learned_legs"]) if "learned_legs" in f.files else 1
        stored_damp = float(f["learned_last_damp"]) if "learned_last_damp" in f.files else 0.0
    legs = stored_legs if legs is None else legs
    damp = stored_damp if damp is None else damp
    stored = (legs, damp) == (stored_legs, stored_damp)
    net = learned.GammaMLP.load(wpath)
    dec = RelayBPDecoder(H, p, data_type="float32", algo="relay", seed=0, device=[0], learned_gamma=net,
                         learned_legs=legs, learned_last_damp=damp, **RELAY)
    dec_t = RelayBPDecoder(H, p, data_type="float32", algo="relay", seed=0, device=[0], learned_gamma=net,
                           learned_legs=legs, learned_last_damp=damp, **RELAY)
    dec_t.requeue_below = args.requeue_below
    ref = dec.decode(syn)
    m = measure(dec_t, syn, args.steps, args.warmup, args.depth)
    seq = measure(dec, syn, max(3, args.steps // 4), 1, 1)
    out = m["out"]
    errors = out["errors"].cpu().numpy()
    conv = out["found_count"].cpu().numpy() >= RELAY["stop_nconv"]
    O = codes.synthetic_observables(H.shape[1], seed=0).astype(np.int64)
    logical = (~conv) | ((O @ (errors ^ err).T.astype(np.int64)) % 2).any(0)
    le = wilson(int(logical.sum()), errors.shape[0])
    rate = m["rate"]
    seq_ms = sum(seq["kernel_ms"]) / len(seq["kernel_ms"])
    roof = roofline(int(out["iterations"].long().sum()), bytes_per_iteration(dec.C, dec.V, dec.E), m["kernel_ms"],
                    m["elapsed"], args.steps, seq_ms, args.depth, _traffic("learned") if stored else None,
                    "relay_decode_kernel<1024, 15, true> (learned strengths)")
    big = ler_eval(dec, H, p)
    return {"shots_per_s": round(rate, 1), "kernel_ms": round(sum(m["kernel_ms"]) / len(m["kernel_ms"]), 4),
            "ler_65536": big,
            "ler_within_plain_ci": bool(plain_ler[1] <= big["ler"][0] <= plain_ler[2]),
            "learned_legs": legs, "learned_last_damp": damp, "speedup_vs_plain": round(rate / plain_rate, 3),
            "sequential_shots_per_s": round(seq["rate"], 1),
            "sequential_kernel_ms": round(sum(seq["kernel_ms"]) / len(seq["kernel_ms"]), 4),
            "mean_iterations": round(float(out["iterations"].float().mean()), 3),
            "synthetic_observable_ler": [round(x, 5) for x in le],
            "matches_decode": same_outputs(m["slots"] + seq["slots"], ref),
            "weights": str(wpath.relative_to(ROOT)), "roofline": roof}


def code_capacity_bb72(args) -> dict:
    """The [[72,12,6]] bivariate bicycle code at code capacity: 1024 shots at
    priors 0.03 (errors from seed 0), the same relay parameters.  Its per-shot
    state is a few KB, so HBM is not the bound: besides shots/s this reports
    the LDS bytes per shot-iteration the design moves (20 B per edge slot over
    both passes, 20 B per variable, 16 B per check; no padding counted) as a
    rate, the per-shot decode latency from the device clock (one launch, no
    re-queue), and the logical error rate over the code's 12 logical X
    operators on 16,384 fresh shots (failure: no solution, or an observable
    flipped; pelagos_benchmark.py:1666-1669)."""
    import numpy as np
    import torch

    from pelagos_amd import RelayBPDecoder, codes
    from pelagos_amd.observables import ObservableDecoderRunner

    H = codes.bb_72_12_6()
    L = codes.logical_operators(H.toarray(), codes.bb_72_12_6_hz())
    p = np.full(H.shape[1], 0.03)
    C, V, E = int(H.shape[0]), int(H.shape[1]), int(H.nnz)
    err = codes.sample_errors(p, 1024, seed=0)
    syn = torch.from_numpy(np.ascontiguousarray(codes.syndromes_of(H, err))).cuda()

    def make():
        return RelayBPDecoder(H, p, data_type="float32", algo="relay", seed=0, device=[0], **RELAY)

    dec, dec_t, dec_l = make(), make(), make()
    dec_t.requeue_below = args.requeue_below
    ref = dec.decode(syn)
    m = measure(dec_t, syn, args.steps, args.warmup, args.depth)
    seq = measure(dec, syn, max(3, args.steps // 4), 1, 1)
    out = m["out"]
    shot_iters = int(out["iterations"].long().sum())
    kern = sum(m["kernel_ms"]) / len(m["kernel_ms"])
    seq_ms = sum(seq["kernel_ms"]) / len(seq["kernel_ms"])
    lds_b = 20 * E + 20 * V + 16 * C
    dec_l.requeue_below = 0
    dec_l.record_shot_clock = True
    dec_l.decode(syn)
    torch.cuda.synchronize()
    clk = dec_l.shot_clock.cpu().numpy()
    us = (clk[:, 1] - clk[:, 0]) / 100.0  # the device wall clock runs at 100 MHz
    ler = ObservableDecoderRunner(dec, L).logical_error_rate(codes.sample_errors(p, 16384, seed=1))
    return {"code": "[[72,12,6]] BB, H_X 36x72, priors 0.03", "shots": 1024,
            "shots_per_s": round(m["rate"], 1), "kernel_ms": round(kern, 4),
            "sequential_shots_per_s": round(seq["rate"], 1), "sequential_kernel_ms": round(seq_ms, 4),
            "mean_iterations": round(shot_iters / 1024, 3),
            "lds_bytes_per_iteration": lds_b,
            "lds_gb_per_s": round(lds_b * shot_iters / (kern * 1e-3) / 1e9, 1),
            "hbm_algorithmic_gb_per_s": round(bytes_per_iteration(C, V, E) * shot_iters / (kern * 1e-3) / 1e9, 1),
            "shot_latency_us": {"p50": round(float(np.percentile(us, 50)), 2),
                                "p99": round(float(np.percentile(us, 99)), 2), "max": round(float(us.max()), 2)},
            "ler": {k: ler[k] for k in ("shots", "failures", "observable_errors", "unconverged",
                                         "logical_error_rate", "wilson95", "mean_iterations")},
            "matches_decode": same_outputs(m["slots"] + seq["slots"], ref)}


def ler_study(batches: int) -> int:
    """Paired logical error rates of the plain form, the shipped learned form
    and the damped fourth leg (shift x 0.5) on the same ``batches`` x 4096
    fresh shots of the 936 x 8785 workload: errors drawn on the GPU from the
    priors (batch k seeded 5000 + k), syndromes and observable flips computed
    exactly in fp64.  A shot fails when it did not converge or an observable
    flipped.  Prints one JSON line: per form the failures, their Wilson 95%
    interval, the mean iterations, and how many failures it shares with the
    plain form."""
    import numpy as np
    import torch

    from pelagos_amd import RelayBPDecoder, codes, learned

    H, p = codes.banded_circuit_code(**WORKLOAD)
    B, V = 4096, H.shape[1]
    dev = torch.device("cuda:0")
    Hd = torch.tensor(H.toarray(), dtype=torch.float64, device=dev)
    Od = torch.tensor(codes.synthetic_observables(V, seed=0), dtype=torch.float64, device=dev)
    pt = torch.tensor(p, dtype=torch.float32, device=dev)
    wpath = learned.default_weights_path()
    with np.load(wpath, allow_pickle=False) as f:
        legs = int(f["learned_legs"]) if "learned_legs" in f.files else 1
        damp = float(f["learned_last_damp"]) if "learned_last_damp" in f.files else 0.0
    net = learned.GammaMLP.load(wpath)
    forms = [("plain", {}), ("learned", dict(learned_gamma=net, learned_legs=legs, learned_last_damp=damp)),
             ("learned_damped_fourth_leg", dict(learned_gamma=net, learned_legs=4, learned_last_damp=0.5))]
    res, fails = {}, {}
    for name, kw in forms:
        dec = RelayBPDecoder(H, p, data_type="float32", algo="relay", seed=0, device=[0], **RELAY, **kw)
        g = torch.Generator(device=dev)
        bad_all, iters = [], 0
        for k in range(batches):
            g.manual_seed(5000 + k)
            e = (torch.rand(B, V, generator=g, device=dev) < pt).to(torch.uint8)
            syn = torch.remainder(e.double() @ Hd.T, 2).to(torch.uint8).contiguous()
            if k == 0 and name == "plain" and not np.array_equal(syn.cpu().numpy(),
                                                                 codes.syndromes_of(H, e.cpu().numpy())):
                raise RuntimeError("device syndromes differ from H e mod 2")
            out = dec.decode(syn)
            conv = out["found_count"] >= dec.stop_nconv
            flipped = torch.remainder((out["errors"] ^ e).double() @ Od.T, 2).any(1)
            bad_all += (torch.nonzero(~conv | flipped).flatten() + k * B).tolist()
            iters += int(out["iterations"].long().sum())
        n = batches * B
        fails[name] = set(bad_all)
        res[name] = {"shots": n, "failures": len(bad_all), "ler": [round(x, 8) for x in wilson(len(bad_all), n)],
                     "mean_iterations": round(iters / n, 3)}
        print(f"[bench] ler study {name}: {len(bad_all)} / {n}", file=sys.stderr, flush=True)
    for name in res:
        res[name]["shared_with_plain"] = len(fails[name] & fails["plain"])
    print(json.dumps({"ler_study": res, "learned_legs": legs, "learned_last_damp": damp}), flush=True)
    return 0


THREE_CODES = {
    # name: (shots, description)
    "bb72": (8192, "[[72,12,6]] BB code, uniform priors"),
    "banded936": (4096, "936x8785 banded"),
    "banded1872": (1024, "1872x17570 banded"),
}


def _three_codes(prior_scale: float):
    """The three codes with priors scaled to mean prior_scale."""
    import numpy as np

    from pelagos_amd import codes

    out = {}
    out["bb72"] = (codes.bb_72_12_6(), np.full(72, prior_scale))
    H, p = codes.banded_circuit_code(**WORKLOAD)
    out["banded936"] = (H, np.minimum(p * (prior_scale / float(p.mean())), 0.45).astype(np.float32))
    H2, p2 = codes.banded_circuit_code(n_blocks=26, V=17570, seed=0)
    out["banded1872"] = (H2, np.minimum(p2 * (prior_scale / float(p2.mean())), 0.45).astype(np.float32))
    return out


def batched_graph(scales=(0.001, 0.003, 0.01)) -> dict:
    """One captured graph launch decoding the three codes' batches together, at
    each prior scale: shots/s over all three codes (HIP events around the
    replay)."""
    import torch

    from pelagos_amd import RelayBPDecoder, codes
    from pelagos_amd.batching import MultiCodeBatch

    res = []
    for sc in scales:
        cs = _three_codes(sc)
        decs = {k: RelayBPDecoder(H, p, data_type="float32", algo="relay", seed=0, device=[0], **RELAY)
                for k, (H, p) in cs.items()}
        shots = {k: THREE_CODES[k][0] for k in cs}
        mb = MultiCodeBatch(decs, shots)
        for k, (H, p) in cs.items():
            e = codes.sample_errors(p, shots[k], seed=3)
            mb.syndromes[k].copy_(torch.from_numpy(codes.syndromes_of(H, e)))
        mb.capture()
        mb.replay()  # warm-up
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 3
        e0.record()
        for _ in range(reps):
            mb.replay()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        conv = {k: float((mb.out[k]["found_count"] >= RELAY["stop_nconv"]).float().mean()) for k in cs}
        total = sum(shots.values())
        res.append({"prior_scale": sc, "ms_per_launch": round(ms, 3), "shots": total,
                    "shots_per_s": round(total / (ms * 1e-3), 1),
                    "converged_frac": {k: round(v, 5) for k, v in conv.items()}})
        del mb, decs
        torch.cuda.empty_cache()
    return {"codes": {k: {"shots": v[0], "code": v[1]} for k, v in THREE_CODES.items()}, "launches": res}


def serving_runs(seconds: float, prior_scale: float = 0.003,
                 rates=(2_000, 10_000, 50_000, 100_000, 200_000)) -> dict:
    """The continuous-batching server on the three codes (priors scaled to mean
    prior_scale, requests of 1-256 shots, code chosen uniformly, Poisson
    arrivals)."""
    import numpy as np
    import torch

    from pelagos_amd import RelayBPDecoder, codes, learned
    from pelagos_amd.serving import DecodeServer, poisson_trace, run_trace

    cs = _three_codes(prior_scale)
    net = None
    wpath = learned.default_weights_path()
    decs = {}
    for k, (H, p) in cs.items():
        kw = {}
        if k == "banded936" and wpath.exists():
            net = net or learned.GammaMLP.load(wpath)
            with np.load(wpath, allow_pickle=False) as f:
                legs = int(f["learned_legs"]) if "learned_legs" in f.files else 1
                damp = float(f["learned_last_damp"]) if "learned_last_damp" in f.files else 0.0
            kw = dict(learned_gamma=net, learned_legs=legs, learned_last_damp=damp)
        decs[k] = RelayBPDecoder(H, p, data_type="float32", algo="relay", seed=0, device=[0], **RELAY, **kw)
    pools = {}
    for k, (H, p) in cs.items():
        e = codes.sample_errors(p, 4096, seed=4)
        pools[k] = torch.from_numpy(codes.syndromes_of(H, e)).cuda()
    server = DecodeServer(decs, max_batch=2048, slots=4, hbm_budget=64 << 30)
    out = []
    for i, rate in enumerate(rates):
        trace = poisson_trace(rate, seconds, list(cs), seed=10 + i)
        st = run_trace(server, trace, pools)
        st["rate"] = rate
        out.append(st)
        print(f"[bench] serving p={prior_scale} {rate}/s: p50 {st['p50_ms']} ms p99 {st['p99_ms']} ms "
              f"{st['shots_per_s']} shots/s", file=sys.stderr, flush=True)
    return {"codes": list(cs), "prior_scale": prior_scale, "seconds_per_rate": seconds, "request_shots": "1-256 uniform",
            "learned_form_on": "banded936" if net is not None else None, "hbm_budget_gb": 64, "slots": 4,
            "rates": out}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--shots", type=int, default=4096)
    ap.add_argument("--depth", type=int, default=3, help="decode launches in flight in the timed steps")
    ap.add_argument("--requeue-below", type=int, default=0,
                    help="re-queue threshold of the timed steps' plan (0: off; the library default is 16)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-strength form")
    ap.add_argument("--no-damped", action="store_true",
                    help="skip the learned variant with a fourth, half-shifted learned leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-latency runs")
    ap.add_argument("--no-batched", action="store_true", help="skip the three-code graph launch")
    ap.add_argument("--no-bb72", action="store_true", help="skip the [[72,12,6]] code-capacity runs")
    ap.add_argument("--ler-study", type=int, default=0, metavar="BATCHES",
                    help="only compare the error rates of the plain, learned and damped forms on the same "
                         "BATCHES x 4096 fresh shots")
    ap.add_argument("--serving-seconds", type=float, default=10.0,
                    help="length of each serving run (one per arrival rate)")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (3 times), the known-size "
                         "reference that scales the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--traffic-from", nargs=4, metavar=("FETCH", "WRITE", "COPY_FETCH", "COPY_WRITE"),
                    help="only turn four rocprofv3 --pmc csv output folders (FETCH_SIZE and WRITE_SIZE "
                         "of a 1-step run of this benchmark, then of --copy-probe) into the HBM-traffic "
                         "record read by the roofline object")
    args = ap.parse_args()
    if args.ler_study:
        return ler_study(args.ler_study)
    if args.traffic_from:
        return traffic_record(*args.traffic_from)
    if args.copy_probe:
        import torch

        n = args.copy_probe << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n, "copies": 3}), flush=True)
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this decoder runs on one GPU per process; shots shard by "
                          "batch across processes, and only --gpus 1 is measured here"}))
        return 2

    import numpy as np
    import torch

    from pelagos_amd import RelayBPDecoder, codes, utils

    dev = torch.device("cuda", 0)
    H, p = codes.banded_circuit_code(**WORKLOAD)
    err = codes.sample_errors(p, args.shots, seed=0)
    syn_np = codes.syndromes_of(H, err)
    dec = RelayBPDecoder(H, p, data_type="float32", algo="relay", seed=0, device=[0],
                         block_size=args.block or None, **RELAY)
    dec_t = RelayBPDecoder(H, p, data_type="float32", algo="relay", seed=0, device=[0],
                           block_size=args.block or None, **RELAY)
    dec_t.requeue_below = args.requeue_below
    C, V, E = dec.C, dec.V, dec.E
    syn = torch.from_numpy(syn_np).to(dev).contiguous()
    plan = dec.plan(args.shots)
    print(f"[bench] banded {C}x{V} E={E} shots={args.shots} block={plan.block} grid={plan.grid} "
          f"placement={plan.placement} lds={plan.lds_bytes} depth={args.depth} requeue_below={args.requeue_below}",
          file=sys.stderr, flush=True)

    ref = dec.decode(syn)
    m = measure(dec_t, syn, args.steps, args.warmup, args.depth)  # the timed region
    seq = measure(dec, syn, max(3, args.steps // 4), 1, 1)        # one launch at a time
    elapsed = m["elapsed"]
    kern_ms = m["kernel_ms"]
    kern_avg_ms = sum(kern_ms) / max(len(kern_ms), 1)
    seq_ms = sum(seq["kernel_ms"]) / max(len(seq["kernel_ms"]), 1)
    out = m["out"]
    matches = same_outputs(m["slots"] + seq["slots"], ref)

    iters = out["iterations"].cpu().numpy().astype(np.int64)
    found = out["found_count"].cpu().numpy()
    errors = out["errors"].cpu().numpy()
    total_iters = int(iters.sum())
    b_iter = bytes_per_iteration(C, V, E)
    roof_plain = roofline(total_iters, b_iter, kern_ms, elapsed, args.steps, seq_ms, args.depth, _traffic("plain"),
                          "relay_decode_kernel<1024, 15, false> (uniform-draw legs)")
    shots_per_s = m["rate"]

    # Quality: frame errors and a seeded synthetic-observable logical error rate.
    valid = (codes.syndromes_of(H, errors) == syn_np).all(1)
    converged = found >= RELAY["stop_nconv"]
    frame_err = (~converged) | (errors != err).any(1)
    O = codes.synthetic_observables(V, seed=0)
    resid = errors ^ err
    logical = (~converged) | ((O.astype(np.int64) @ resid.T.astype(np.int64)) % 2).any(0)
    fe = wilson(int(frame_err.sum()), args.shots)
    le = wilson(int(logical.sum()), args.shots)

    result = {
        "metric": "decoded shots/s at matched LER (learned vs plain gamma) + p99 latency, 1 GPU",
        "value": round(shots_per_s, 1),
        "value_form": "plain",
        "unit": "shots/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"banded circuit-shaped H {C}x{V} (E={E}, 13 blocks of 72 checks, column "
                        f"weight 3-6), log-uniform priors [1e-4,1e-2], {args.shots} shots, seed 0",
            "shots": args.shots,
            "pre_iter": RELAY["pre_iter"], "num_sets": RELAY["num_sets"],
            "set_max_iter": RELAY["set_max_iter"], "stop_nconv": RELAY["stop_nconv"],
            "gamma0": RELAY["gamma0"], "gamma_dist_interval": list(RELAY["gamma_dist_interval"]),
            "block": plan.block, "grid": plan.grid, "lds_bytes": plan.lds_bytes,
            "launches_in_flight": args.depth, "requeue_below": args.requeue_below,
        },
        "roofline": roof_plain,
        "quality": {
            "mean_iterations": round(float(iters.mean()), 3),
            "converged_frac": round(float(converged.mean()), 5),
            "syndrome_valid_frac": round(float(valid.mean()), 5),
            "frame_error_rate": [round(x, 5) for x in fe],
            "synthetic_observable_ler": [round(x, 5) for x in le],
        },
    }
    result["plain"] = {"shots_per_s": round(shots_per_s, 1), "kernel_ms": round(kern_avg_ms, 4),
                       "sequential_shots_per_s": round(seq["rate"], 1), "sequential_kernel_ms": round(seq_ms, 4),
                       "matches_decode": matches,
                       "synthetic_observable_ler": [round(x, 5) for x in le]}
    if not args.no_learned:
        big = ler_eval(dec, H, p)
        result["plain"]["ler_65536"] = big
        result["learned"] = learned_form(H, p, err, syn, syn_np, args, shots_per_s, big["ler"])
        lr = result["learned"]
        if lr.get("ler_within_plain_ci") and lr.get("shots_per_s", 0) > shots_per_s:
            result["value"] = lr["shots_per_s"]
            result["value_form"] = "learned"
            result["ms_per_step"] = round(args.shots / lr["shots_per_s"] * 1e3, 3)
            result["roofline"] = lr.pop("roofline")
            result["plain"]["roofline"] = roof_plain
        if not args.no_damped and "skipped" not in lr:
            # Not the default: a fourth learned leg at half shift trades a few
            # failures per 1e5 shots for fewer iterations.
            dv = learned_form(H, p, err, syn, syn_np, args, shots_per_s, big["ler"], legs=4, damp=0.5)
            dv.pop("roofline", None)
            dv["default"] = False
            result["learned_damped_fourth_leg"] = dv
    copy_gbs = copy_bandwidth()
    for r in (result.get("roofline"), result["plain"].get("roofline"), result.get("learned", {}).get("roofline")):
        if r:
            r["measured_copy_gbs"] = copy_gbs
            r["frac_of_measured_copy"] = round(r["achieved"] / copy_gbs, 4)
    if not args.no_bb72:
        result["code_capacity_bb72"] = code_capacity_bb72(args)
    if not args.no_batched:
        result["batched"] = batched_graph()
    if not args.no_serving:
        result["serving"] = {f"prior_{sc}": serving_runs(args.serving_seconds, sc) for sc in (0.001, 0.003)}
        result["p99_latency_ms"] = {f"prior_{sc}": {str(r["rate"]): r["p99_ms"] for r in sv["rates"]}
                                    for sc, sv in zip((0.001, 0.003), result["serving"].values())}
    if not args.no_cpu:
        lam = utils.compute_log_prior_ratios(p)
        gtab = dec.gamma_table(args.shots).cpu().numpy()
        cb, ref, n2 = cpu_baseline(dec, syn_np, lam, gtab)
        result["cpu_baseline"] = cb
        same = all(np.array_equal(out[k][:n2].cpu().numpy(), ref[k])
                   for k in ("errors", "iterations", "found_count", "weights"))
        result["cpu_baseline"]["gpu_matches_on_sample"] = bool(same)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
