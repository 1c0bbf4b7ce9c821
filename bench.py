"""Decoded shots/s of Relay-BP-S on one MI355X (936 x 8785 banded circuit-shaped code),
plain and with learned strengths, plus serving latency.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--shots 4096] [--jobs 50]
                    [--no-cpu] [--no-learned] [--no-serving] [--no-batched] [--no-bb72]

A step decodes the whole 4096-shot batch (syndromes already resident in HBM)
into its own output buffers.  The timed steps run as persistent launches of
``--jobs`` steps each (``pel_decode_jobs``: one launch decodes that many
batches, its workgroups moving to the next batch as soon as the current one
has no shot left, so a batch's slow last shots overlap the next batches),
issued back to back on one stream.  Every step's results are checked against
a plain ``decode()`` of the batch.

Prints ONE JSON line with the throughput of the plain form (uniform-draw legs)
and of the learned form (trained weights) on the same shots; ``value`` is the
learned form's rate when its logical error rate, on a paired sample at mean
prior 0.003 where failures are seen, lies in the plain form's Wilson 95%
interval and the paired discordance test does not reject equality, else the
plain rate.  Beside it, the same network with one learned leg fewer (one more
uniform exploration leg): its rate and its paired error rate on the same
shots.  Also: the decode kernel's roofline object (algorithmic bytes per
launch over the launch time, with the measured HBM traffic and the shader
counters of profiles/ beside it), one launch at a time, the [[72,12,6]] code
at code capacity with its CPU baseline, one launch over the three codes at
three prior scales, the server's latency at five Poisson arrival rates, and
the CPU oracle timed on a sample of the same shots on this host.
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
LDS_PEAK_B_PER_CLK_CU = 128
N_XCD = 8  # MI355X: 8 XCDs of 32 CUs
WORKLOAD = dict(n_blocks=13, checks_per_block=72, V=8785, seed=0)
RELAY = dict(pre_iter=80, num_sets=100, set_max_iter=60, gamma0=0.65,
             gamma_dist_interval=(-0.24, 0.66), stop_nconv=5)
PROFILES = ROOT / "profiles"
TRAFFIC_FILE = PROFILES / "engine_banded936x8785_traffic.json"
COUNTERS_FILE = PROFILES / "engine_banded936x8785_counters.json"


def wilson(k: int, n: int, z: float = 1.959964):
    if n == 0:
        return (0.0, 0.0, 1.0)
    p = k / n
    den = 1 + z * z / n
    c = (p + z * z / (2 * n)) / den
    h = z * math.sqrt(p * (1 - p) / n + z * z / (4 * n * n)) / den
    return (p, max(0.0, c - h), min(1.0, c + h))


def mcnemar_p(b: int, c: int) -> float:
    """Exact two-sided McNemar p-value of b against c discordant pairs."""
    n = b + c
    if n == 0:
        return 1.0
    k = min(b, c)
    tail = sum(math.comb(n, i) for i in range(k + 1)) / 2.0 ** n
    return min(1.0, 2.0 * tail)


def bytes_per_iteration(C: int, V: int, E: int) -> int:
    """Algorithmic HBM bytes of one BP iteration of one shot with edge messages in
    HBM (fp32): read mu + write nu (C2V), read nu + write mu (V2C), read+write the
    posterior, bit-packed hard decisions and syndrome."""
    return 16 * E + 8 * V + (V + 7) // 8 + (C + 7) // 8


def lds_bytes_per_iteration(H) -> int:
    """LDS bytes one shot-iteration moves in this design: 12 B per padded edge
    slot of the check pass (the variable state and one check message), 8 B per
    padded slot of the variable pass (code word and message), 20 B per variable
    and 16 B per check.  Rows and variables are in 64-wide slices of equal-
    degree-sorted entries, rows padded to whole chunks of 8."""
    import numpy as np

    H = H.tocsr()
    rdeg = np.sort(np.diff(H.indptr))[::-1]
    vdeg = np.sort(np.bincount(H.indices, minlength=H.shape[1]))[::-1]
    rslots = sum(64 * int(-(-int(rdeg[i]) // 8) * 8) for i in range(0, len(rdeg), 64))
    vslots = sum(64 * int(vdeg[i]) for i in range(0, len(vdeg), 64))
    return 12 * rslots + 8 * vslots + 20 * H.shape[1] + 16 * H.shape[0]


def _json_or_none(path: Path):
    try:
        return json.loads(path.read_text())
    except Exception:
        return None


def cpu_baseline(graph, syn_np, lam, gtab, target_s: float = 10.0, single_sample: int = 256):
    """The C oracle (one shot per OpenMP thread, all host cores) on the same
    shots and gamma table, repeated over a prefix of the batch until about
    target_s of work; plus a single-thread figure on a smaller prefix."""
    import oracle

    threads = os.cpu_count() or 1
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    kw = dict(pre_iter=RELAY["pre_iter"], num_sets=RELAY["num_sets"],
              set_max_iter=RELAY["set_max_iter"], stop_nconv=RELAY["stop_nconv"],
              gamma0=RELAY["gamma0"])
    B = syn_np.shape[0]
    done, reps, ref = 0, 0, None
    t0 = time.perf_counter()
    while True:
        out = oracle.decode(graph.chk_ptr, graph.edge_var, graph.V, syn_np, lam, gtab, nthreads=threads, **kw)
        ref = ref or out
        done += B
        reps += 1
        dt = time.perf_counter() - t0
        if dt >= target_s or reps >= 10000:
            break
    n1 = min(B, single_sample)
    t1 = time.perf_counter()
    oracle.decode(graph.chk_ptr, graph.edge_var, graph.V, syn_np[:n1], lam, gtab, nthreads=1, **kw)
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
    }, ref


# The engine instantiation of each form, by its demangled or mangled name.
ENGINE_FORMS = (("plain", ("engine_kernel<1>", "engine_kernelILi1E")),
                ("learned", ("engine_kernel<32>", "engine_kernelILi32E")))


def _pmc_rows(folder: str):
    import csv
    import glob

    rows = []
    for f in sorted(glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True)):
        with open(f, newline="") as fh:
            rows += list(csv.DictReader(fh))
    return rows


def _per_dispatch(rows, kernel_pred, counter):
    """{dispatch id: summed counter value} over the dispatches of matching kernels."""
    out = {}
    for r in rows:
        if kernel_pred(r["Kernel_Name"]) and r["Counter_Name"] == counter:
            d = r.get("Dispatch_Id") or r.get("Correlation_Id") or str(len(out))
            out[d] = out.get(d, 0.0) + float(r["Counter_Value"])
    return out


def traffic_record(fetch_dir: str, write_dir: str, cfetch_dir: str, cwrite_dir: str) -> int:
    """HBM bytes per engine launch from FETCH_SIZE / WRITE_SIZE (KiB per
    dispatch, averaged over the timed launches of each form), each scaled by
    what the same counter reads for a device copy of known size."""
    copy_bytes = 1024 << 20
    pred_copy = lambda k: "copyBuffer" in k  # noqa: E731
    cp_f = list(_per_dispatch(_pmc_rows(cfetch_dir), pred_copy, "FETCH_SIZE").values())
    cp_w = list(_per_dispatch(_pmc_rows(cwrite_dir), pred_copy, "WRITE_SIZE").values())
    f_scale = copy_bytes / (sum(cp_f) / len(cp_f) * 1024)
    w_scale = copy_bytes / (sum(cp_w) / len(cp_w) * 1024)
    rec = {"kernel": "engine_kernel", "workload": "banded 936x8785, launches of 50 batches of 4096 shots (bench.py --jobs 50)",
           "copy_probe_bytes": copy_bytes, "copy_fetch_kib_raw": sum(cp_f) / len(cp_f),
           "copy_write_kib_raw": sum(cp_w) / len(cp_w), "fetch_scale": round(f_scale, 4),
           "write_scale": round(w_scale, 4)}
    frows, wrows = _pmc_rows(fetch_dir), _pmc_rows(write_dir)
    for form, tags in ENGINE_FORMS:
        pred = (lambda k, tags=tags: any(t in k for t in tags))
        f = list(_per_dispatch(frows, pred, "FETCH_SIZE").values())
        w = list(_per_dispatch(wrows, pred, "WRITE_SIZE").values())
        if not f or not w:
            continue
        rec[form] = {"launches": len(f), "fetch_kib_raw": sum(f) / len(f), "write_kib_raw": sum(w) / len(w),
                     "hbm_bytes_per_launch": int(sum(f) / len(f) * 1024 * f_scale + sum(w) / len(w) * 1024 * w_scale)}
    TRAFFIC_FILE.write_text(json.dumps(rec, indent=1) + "\n")
    print(json.dumps(rec))
    return 0


def counters_record(dirs) -> int:
    """Shader counters per engine launch (averaged over each form's launches):
    SQ_INSTS_VALU, SQ_INSTS_LDS, SQ_LDS_BANK_CONFLICT, SQ_WAVE_CYCLES,
    SQ_BUSY_CYCLES and GRBM_GUI_ACTIVE, from rocprofv3 --pmc output folders."""
    rows = []
    for d in dirs:
        rows += _pmc_rows(d)
    names = sorted({r["Counter_Name"] for r in rows})
    rec = {"kernel": "engine_kernel", "workload": "banded 936x8785, launches of 50 batches of 4096 shots (bench.py --jobs 50)",
           "counters": names}
    for form, tags in ENGINE_FORMS:
        pred = (lambda k, tags=tags: any(t in k for t in tags))
        vals = {}
        for c in names:
            v = list(_per_dispatch(rows, pred, c).values())
            if v:
                vals[c] = sum(v) / len(v)
        if vals:
            rec[form] = vals
    COUNTERS_FILE.write_text(json.dumps(rec, indent=1) + "\n")
    print(json.dumps(rec))
    return 0


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


def same_outputs(out, ref) -> bool:
    import torch

    return all(torch.equal(out[k], ref[k]) for k in ("errors", "weights", "iterations", "found_count"))


def engine_measure(dec, syn, steps: int, warmup: int, jobs: int) -> dict:
    """``steps`` decodes of the batch ``syn`` as launches of ``jobs`` batches each
    (a shorter last launch when jobs does not divide steps), back to back on
    one stream, after ``warmup`` untimed steps, bracketed by device
    synchronisation.  Returns the rate, the per-launch kernel times (HIP events
    recorded by the library around each launch), the shot-iterations of one
    batch, and whether every batch of the last launches equals ``decode()``."""
    import torch

    from pelagos_amd.batching import JobLaunch

    B = int(syn.shape[0])
    jobs = max(1, min(jobs, steps))  # fewer steps than batches per launch: one launch of all of them
    full = JobLaunch([(f"s{i}", dec, B) for i in range(jobs)])
    rem = steps % jobs
    part = JobLaunch([(f"s{i}", dec, B) for i in range(rem)]) if rem else None
    for jl in (full, part):
        if jl is not None:
            for name in jl.names:
                jl.syndromes[name].copy_(syn)
    for _ in range(max(1, -(-warmup // jobs))):
        full.launch()
    torch.cuda.synchronize()
    ev = []
    full.kernel_events = ev
    t0 = time.perf_counter()
    for _ in range(steps // jobs):
        full.launch()
    if part is not None:
        part.kernel_events = ev
        part.launch()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    full.kernel_events = None
    kern = [a.elapsed_time(b) for a, b in ev]
    ref = dec.decode(syn)
    torch.cuda.synchronize()
    matches = all(same_outputs(jl.out[n], ref) for jl in (full, part) if jl is not None for n in jl.names)
    iters = int(ref["iterations"].long().sum())
    return {"rate": B * steps / elapsed, "elapsed": elapsed, "kernel_ms": kern[:steps // jobs],
            "launches": len(kern), "jobs_per_launch": jobs, "batch_iterations": iters, "matches": matches,
            "roles": full.roles()["s0"], "ref": ref}


def one_at_a_time(dec, syn, reps: int) -> dict:
    """decode() of the batch, one launch at a time (the library's default plan,
    re-queue of fast first legs included): HIP-event kernel time per launch."""
    import torch

    dec.decode(syn)
    torch.cuda.synchronize()
    ev = []
    dec.kernel_events = ev
    t0 = time.perf_counter()
    for _ in range(reps):
        dec.decode(syn)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    dec.kernel_events = None
    ms = sum(a.elapsed_time(b) for a, b in ev) / len(ev)
    return {"shots_per_s": round(int(syn.shape[0]) * reps / elapsed, 1), "kernel_ms": round(ms, 4)}


def roofline(form: str, m: dict, b_iter: int, lds_iter: int, n_cu: int, copy_gbs: float) -> dict:
    """The decode kernel's roofline object for one form, from launches that do
    not share the GPU (one stream).

    achieved: algorithmic bytes per launch (B_iter x the launch's shot-
    iterations, messages counted as if in HBM) over the average launch time
    (HIP events).  The design keeps every message in LDS, so this can exceed
    the HBM peak; ``frac_exceeds_1`` says so.  traffic: the HBM bytes per
    launch measured by FETCH_SIZE / WRITE_SIZE (profiles/), with the rate and
    fraction of peak it represents.  The kernel's real limiter is instruction
    issue inside the CU: valu_issue_frac = VALU wave-instructions x 4 cycles
    over (CUs x 4 SIMDs x launch cycles), the launch cycles being
    GRBM_GUI_ACTIVE over the XCDs (it sums their busy cycles), and the LDS
    rate per clock per CU of the design's LDS bytes over those cycles."""
    kern_ms = sum(m["kernel_ms"]) / max(len(m["kernel_ms"]), 1)
    iters_launch = m["batch_iterations"] * m["jobs_per_launch"]
    algo = iters_launch * b_iter
    achieved = algo / (kern_ms * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "frac_exceeds_1": achieved > HBM_PEAK_GBS,
           "achieved_is": "algorithmic bytes (B_iter per shot-iteration, messages counted as if in HBM), "
                          "not traffic",
           "kernel": "engine_kernel<1> (plain)" if form == "plain" else "engine_kernel<32> (learned)",
           "kernel_ms": round(kern_ms, 4), "jobs_per_launch": m["jobs_per_launch"],
           "bytes_per_iteration": b_iter, "iterations_per_launch": iters_launch,
           "measured_copy_gbs": copy_gbs, "frac_of_measured_copy": round(achieved / copy_gbs, 4),
           "traffic": None, "limiter": "VALU and LDS instruction issue inside the CU (every message in LDS)"}
    tr = _json_or_none(TRAFFIC_FILE) or {}
    t = tr.get(form, {}).get("hbm_bytes_per_launch")
    if t:
        out["traffic"] = int(t)
        out["traffic_gbs"] = round(t / (kern_ms * 1e-3) / 1e9, 1)
        out["traffic_frac_of_peak"] = round(t / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)
        out["traffic_source"] = str(TRAFFIC_FILE.relative_to(ROOT))
    ct = (_json_or_none(COUNTERS_FILE) or {}).get(form)
    if ct and "SQ_INSTS_VALU" in ct and "GRBM_GUI_ACTIVE" in ct:
        cyc = ct["GRBM_GUI_ACTIVE"] / N_XCD  # the counter sums the busy cycles of every XCD
        out["valu_issue_frac"] = round(ct["SQ_INSTS_VALU"] * 4 / (n_cu * 4 * cyc), 4)
        out["clock_ghz"] = round(cyc / (kern_ms * 1e-3) / 1e9, 3)
        out["lds_bytes_per_clk_per_cu"] = round(lds_iter * iters_launch / (n_cu * cyc), 2)
        out["lds_frac_of_peak"] = round(lds_iter * iters_launch / (n_cu * cyc) / LDS_PEAK_B_PER_CLK_CU, 4)
        if "SQ_INSTS_LDS" in ct:
            out["lds_insts_per_iteration"] = round(ct["SQ_INSTS_LDS"] / iters_launch, 1)
        out["valu_insts_per_iteration"] = round(ct["SQ_INSTS_VALU"] / iters_launch, 1)
        if "SQ_LDS_BANK_CONFLICT" in ct:
            out["lds_bank_conflict_cycles_per_iteration"] = round(ct["SQ_LDS_BANK_CONFLICT"] / iters_launch, 1)
        out["counters_source"] = str(COUNTERS_FILE.relative_to(ROOT))
    return out


def paired_ler(decs: dict, H, p, batches: int, shots: int, seed0: int) -> dict:
    """Logical error rates of several decoders on the same ``batches`` x
    ``shots`` fresh shots (errors from the priors p, seeds seed0 + k): a shot
    fails when it did not converge or its decoding differs from the true error
    by one of 12 seeded synthetic observables.  Per form: failures, Wilson 95%
    interval, mean iterations; per form but the first: the discordant pairs
    against the first form and the exact McNemar p-value."""
    import numpy as np
    import torch

    from pelagos_amd import codes

    # float32 products of 0/1 matrices are exact (counts stay below 2^24) and run on BLAS
    O = codes.synthetic_observables(H.shape[1], seed=0).astype(np.float32)
    fails = {k: np.zeros(0, bool) for k in decs}
    iters = {k: 0 for k in decs}
    for k in range(batches):
        err = codes.sample_errors(p, shots, seed=seed0 + k)
        syn = torch.from_numpy(codes.syndromes_of(H, err)).to("cuda:0").contiguous()
        for name, dec in decs.items():
            out = dec.decode(syn)
            e = out["errors"].cpu().numpy()
            conv = out["found_count"].cpu().numpy() >= dec.stop_nconv
            bad = (~conv) | ((O @ (e ^ err).T.astype(np.float32)).astype(np.int64) % 2).any(0)
            fails[name] = np.concatenate([fails[name], bad])
            iters[name] += int(out["iterations"].long().sum())
        if (k + 1) % 64 == 0 and k + 1 < batches:
            print(f"[bench] paired error rates: {k + 1} of {batches} batches", file=sys.stderr, flush=True)
    n = batches * shots
    names = list(decs)
    res = {"shots": n}
    for name in names:
        f = fails[name]
        r = {"failures": int(f.sum()), "ler": [round(x, 7) for x in wilson(int(f.sum()), n)],
             "mean_iterations": round(iters[name] / n, 3)}
        if name != names[0]:
            b = int((fails[names[0]] & ~f).sum())
            c = int((~fails[names[0]] & f).sum())
            r["discordant"] = {f"{names[0]}_only": b, f"{name}_only": c}
            r["mcnemar_p"] = round(mcnemar_p(b, c), 4)
        res[name] = r
    return res


def code_capacity_bb72(args) -> dict:
    """The [[72,12,6]] bivariate bicycle code at code capacity: 1024 shots at
    priors 0.03 (errors from seed 0), the same relay parameters.  One launch of
    ``--jobs`` batches runs a shot per wave (sixteen per workgroup).  Besides
    shots/s: the LDS bytes per shot-iteration the design moves as a rate, the
    per-shot decode latency from the device clock (the single-batch decode,
    one 64-thread workgroup per shot), the logical error rate over the code's
    12 logical X operators on 16,384 fresh shots, and the CPU oracle on the
    same 1024 shots."""
    import numpy as np
    import torch

    from pelagos_amd import RelayBPDecoder, codes
    from pelagos_amd.observables import ObservableDecoderRunner

    H = codes.bb_72_12_6()
    L = codes.logical_operators(H.toarray(), codes.bb_72_12_6_hz())
    p = np.full(H.shape[1], 0.03)
    C, V, E = int(H.shape[0]), int(H.shape[1]), int(H.nnz)
    err = codes.sample_errors(p, 1024, seed=0)
    syn_np = np.ascontiguousarray(codes.syndromes_of(H, err))
    syn = torch.from_numpy(syn_np).cuda()
    dec = RelayBPDecoder(H, p, data_type="float32", algo="relay", seed=0, device=[0], **RELAY)
    m = engine_measure(dec, syn, args.steps, args.warmup, args.jobs)
    seq = one_at_a_time(dec, syn, max(3, args.steps // 4))
    shot_iters = m["batch_iterations"]
    kern = sum(m["kernel_ms"]) / len(m["kernel_ms"]) / m["jobs_per_launch"]
    lds_b = 20 * E + 20 * V + 16 * C
    dec_l = RelayBPDecoder(H, p, data_type="float32", algo="relay", seed=0, device=[0], **RELAY)
    dec_l.requeue_below = 0
    dec_l.record_shot_clock = True
    dec_l.decode(syn)
    torch.cuda.synchronize()
    clk = dec_l.shot_clock.cpu().numpy()
    us = (clk[:, 1] - clk[:, 0]) / 100.0  # the device wall clock runs at 100 MHz
    ler = ObservableDecoderRunner(dec, L).logical_error_rate(codes.sample_errors(p, 16384, seed=1))
    res = {"code": "[[72,12,6]] BB, H_X 36x72, priors 0.03", "shots": 1024,
           "shots_per_s": round(m["rate"], 1), "kernel_ms_per_batch": round(kern, 4),
           "role": m["roles"], "jobs_per_launch": m["jobs_per_launch"],
           "one_at_a_time": seq, "mean_iterations": round(shot_iters / 1024, 3),
           "lds_bytes_per_iteration": lds_b,
           "lds_gb_per_s": round(lds_b * shot_iters / (kern * 1e-3) / 1e9, 1),
           "hbm_algorithmic_gb_per_s": round(bytes_per_iteration(C, V, E) * shot_iters / (kern * 1e-3) / 1e9, 1),
           "shot_latency_us": {"p50": round(float(np.percentile(us, 50)), 2),
                               "p99": round(float(np.percentile(us, 99)), 2), "max": round(float(us.max()), 2)},
           "ler": {k: ler[k] for k in ("shots", "failures", "observable_errors", "unconverged",
                                        "logical_error_rate", "wilson95", "mean_iterations")},
           "matches_decode": m["matches"]}
    if not args.no_cpu:
        from pelagos_amd import utils

        lam = utils.compute_log_prior_ratios(p)
        cb, ref = cpu_baseline(dec.graph, syn_np, lam, dec.gamma_table(1024).cpu().numpy(), target_s=5.0)
        cb["gpu_matches_on_sample"] = all(np.array_equal(m["ref"][k].cpu().numpy(), ref[k])
                                          for k in ("errors", "iterations", "found_count", "weights"))
        cb["gpu_speedup"] = round(m["rate"] / cb["value"], 1)
        res["cpu_baseline"] = cb
    return res


def batched_three_codes(scales=(0.001, 0.003, 0.01)) -> dict:
    """One launch over the three codes' batches -- the [[72,12,6]] code (8192
    shots, a shot per wave), the 936 x 8785 code (4096) and the 1872 x 17570
    code (1024, its state spilled) -- at each prior scale: shots/s over all
    three (HIP events around the launch), each code's convergence, and that
    each code's results equal its own decode()."""
    import torch

    from pelagos_amd import RelayBPDecoder, codes
    from pelagos_amd.batching import MultiCodeBatch

    shots = {"bb72": 8192, "banded936": 4096, "banded1872": 1024}
    res = []
    roles = None
    for sc in scales:
        cs = codes.three_code_workload(sc)
        decs = {k: RelayBPDecoder(H, p, data_type="float32", algo="relay", seed=0, device=[0], **RELAY)
                for k, (H, p) in cs.items()}
        mb = MultiCodeBatch(decs, shots)
        roles = mb.roles()
        syn = {}
        for k, (H, p) in cs.items():
            syn[k] = torch.from_numpy(codes.syndromes_of(H, codes.sample_errors(p, shots[k], seed=3))).cuda()
            mb.syndromes[k].copy_(syn[k])
        mb.launch()  # warm-up
        torch.cuda.synchronize()
        ev = []
        mb.kernel_events = ev
        reps = 3
        for _ in range(reps):
            mb.launch()
        torch.cuda.synchronize()
        mb.kernel_events = None
        ms = sum(a.elapsed_time(b) for a, b in ev) / reps
        conv = {k: float((mb.out[k]["found_count"] >= RELAY["stop_nconv"]).float().mean()) for k in cs}
        match = all(same_outputs(mb.out[k], decs[k].decode(syn[k].contiguous())) for k in cs)
        total = sum(shots.values())
        res.append({"prior_scale": sc, "ms_per_launch": round(ms, 3), "shots": total,
                    "shots_per_s": round(total / (ms * 1e-3), 1), "matches_decode": match,
                    "converged_frac": {k: round(v, 5) for k, v in conv.items()}})
        print(f"[bench] three codes p={sc}: {ms:.2f} ms per launch", file=sys.stderr, flush=True)
        del mb, decs
        torch.cuda.empty_cache()
    return {"codes": {k: {"shots": v} for k, v in shots.items()}, "roles": roles, "launches": res}


def serving_runs(seconds: float, prior_scale: float,
                 rates=(2_000, 10_000, 50_000, 100_000, 200_000, 400_000)) -> dict:
    """The server on the three codes (priors scaled to mean prior_scale,
    requests of 1-256 shots, code chosen uniformly, Poisson arrivals, learned
    form on the 936 x 8785 code): one persistent serving launch that admits
    each request's shots as they arrive, with a bounded queue that refuses work
    beyond it.  ``saturation_shots_per_s`` is the highest served rate among the
    offered rates at which under 1 % of the offered shots were refused."""
    import numpy as np
    import torch

    from pelagos_amd import RelayBPDecoder, codes, learned
    from pelagos_amd.serving import DecodeServer, poisson_trace, run_trace

    cs = codes.three_code_workload(prior_scale)
    wpath = learned.default_weights_path()
    decs = {}
    for k, (H, p) in cs.items():
        kw = {}
        if k == "banded936" and wpath.exists():
            with np.load(wpath, allow_pickle=False) as f:
                legs = int(f["learned_legs"])
            kw = dict(learned_gamma=learned.GammaMLP.load(wpath), learned_legs=legs)
        decs[k] = RelayBPDecoder(H, p, data_type="float32", algo="relay", seed=0, device=[0], **RELAY, **kw)
    pools = {k: torch.from_numpy(codes.syndromes_of(H, codes.sample_errors(p, 4096, seed=4))).cuda()
             for k, (H, p) in cs.items()}
    torch.cuda.synchronize()
    out = []
    with DecodeServer(decs, max_batch=2048, max_queued_shots=16384, idle_ms=2.0) as server:
        roles = server.roles()
        for i, rate in enumerate(rates):
            trace = poisson_trace(rate, seconds, list(cs), seed=10 + i)
            l0 = server.launches
            st = run_trace(server, trace, pools, policy="reject")
            st["rate"] = rate
            st["offered_shots"] = int(sum(b for _, _, b in trace))
            st["engine_launches"] = server.launches - l0
            out.append(st)
            print(f"[bench] serving p={prior_scale} {rate}/s: p50 {st['p50_ms']} ms p99 {st['p99_ms']} ms "
                  f"{st['shots_per_s']} shots/s, refused {st['rejected_requests']}, "
                  f"launches {st['engine_launches']}", file=sys.stderr, flush=True)
    ok = [st["shots_per_s"] for st in out if st["rejected_shots"] < 0.01 * max(1, st["offered_shots"])]
    return {"codes": list(cs), "prior_scale": prior_scale, "seconds_per_rate": seconds,
            "request_shots": "1-256 uniform", "learned_form_on": "banded936" if wpath.exists() else None,
            "engine": "one persistent serving launch, shots admitted as requests arrive (pel_serve_launch)",
            "roles": roles, "max_queued_shots": 16384, "policy": "reject",
            "saturation_shots_per_s": max(ok) if ok else None, "rates": out}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--shots", type=int, default=4096)
    ap.add_argument("--jobs", type=int, default=50, help="batches (steps) per persistent launch")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baselines")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-strength form")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-latency runs")
    ap.add_argument("--no-batched", action="store_true", help="skip the three-code launch")
    ap.add_argument("--no-bb72", action="store_true", help="skip the [[72,12,6]] code-capacity runs")
    ap.add_argument("--no-fewer-legs", action="store_true",
                    help="skip the network with one learned leg fewer (its launches share the learned kernel)")
    ap.add_argument("--ler-batches", type=int, default=64,
                    help="4096-shot batches of the paired error-rate comparison")
    ap.add_argument("--serving-seconds", type=float, default=5.0,
                    help="length of each serving run (one per arrival rate)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (3 times), the known-size "
                         "reference that scales the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--traffic-from", nargs=4, metavar=("FETCH", "WRITE", "COPY_FETCH", "COPY_WRITE"),
                    help="only turn four rocprofv3 --pmc csv output folders (FETCH_SIZE and WRITE_SIZE "
                         "of a short run of this benchmark, then of --copy-probe) into the HBM-traffic "
                         "record read by the roofline object")
    ap.add_argument("--counters-from", nargs="+", metavar="DIR",
                    help="only turn rocprofv3 --pmc csv output folders of a short run of this benchmark "
                         "into the shader-counter record read by the roofline object")
    args = ap.parse_args()
    if args.traffic_from:
        return traffic_record(*args.traffic_from)
    if args.counters_from:
        return counters_record(args.counters_from)
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
    if not 1 <= args.jobs <= 64:
        print(json.dumps({"error": "--jobs must lie in 1..64 (PEL_MAX_JOBS)"}))
        return 2

    import numpy as np
    import torch

    from pelagos_amd import RelayBPDecoder, codes, learned, utils

    H, p = codes.banded_circuit_code(**WORKLOAD)
    err = codes.sample_errors(p, args.shots, seed=0)
    syn_np = codes.syndromes_of(H, err)
    syn = torch.from_numpy(syn_np).to("cuda:0").contiguous()
    dec = RelayBPDecoder(H, p, data_type="float32", algo="relay", seed=0, device=[0], **RELAY)
    C, V, E = dec.C, dec.V, dec.E
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    b_iter = bytes_per_iteration(C, V, E)
    lds_iter = lds_bytes_per_iteration(H)
    print(f"[bench] banded {C}x{V} E={E} shots={args.shots} jobs/launch={args.jobs} steps={args.steps}",
          file=sys.stderr, flush=True)

    m = engine_measure(dec, syn, args.steps, args.warmup, args.jobs)     # the timed region (plain)
    print(f"[bench] plain {m['rate']:.0f} shots/s", file=sys.stderr, flush=True)
    seq = one_at_a_time(dec, syn, max(3, args.steps // 8))
    ref = m["ref"]
    iters = ref["iterations"].cpu().numpy().astype(np.int64)
    found = ref["found_count"].cpu().numpy()
    errors = ref["errors"].cpu().numpy()
    valid = (codes.syndromes_of(H, errors) == syn_np).all(1)
    converged = found >= RELAY["stop_nconv"]
    O = codes.synthetic_observables(V, seed=0)
    logical = (~converged) | ((O.astype(np.int64) @ (errors ^ err).T.astype(np.int64)) % 2).any(0)
    copy_gbs = copy_bandwidth()
    roof_plain = roofline("plain", m, b_iter, lds_iter, n_cu, copy_gbs)

    result = {
        "metric": "decoded shots/s at matched LER (learned vs plain gamma) + p99 latency, 1 GPU",
        "value": round(m["rate"], 1),
        "value_form": "plain",
        "unit": "shots/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(m["elapsed"] / args.steps * 1e3, 3),
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
            "launch": f"pel_decode_jobs, {m['jobs_per_launch']} batches per launch, one stream",
            "role": m["roles"],
        },
        "roofline": roof_plain,
        "quality": {
            "mean_iterations": round(float(iters.mean()), 3),
            "converged_frac": round(float(converged.mean()), 5),
            "syndrome_valid_frac": round(float(valid.mean()), 5),
            "synthetic_observable_ler": [round(x, 5) for x in wilson(int(logical.sum()), args.shots)],
        },
        "plain": {"shots_per_s": round(m["rate"], 1),
                  "kernel_ms_per_launch": round(sum(m["kernel_ms"]) / len(m["kernel_ms"]), 4),
                  "one_at_a_time": seq, "matches_decode": m["matches"]},
    }
    if not args.no_learned and learned.default_weights_path().exists():
        wpath = learned.default_weights_path()
        with np.load(wpath, allow_pickle=False) as f:
            legs = int(f["learned_legs"])
            damp = float(f["learned_last_damp"]) if "learned_last_damp" in f.files else 0.0
        net = learned.GammaMLP.load(wpath)
        dl = RelayBPDecoder(H, p, data_type="float32", algo="relay", seed=0, device=[0], learned_gamma=net,
                            learned_legs=legs, learned_last_damp=damp, **RELAY)
        ml = engine_measure(dl, syn, args.steps, args.warmup, args.jobs)
        print(f"[bench] learned {ml['rate']:.0f} shots/s", file=sys.stderr, flush=True)
        seql = one_at_a_time(dl, syn, max(3, args.steps // 8))
        roof_l = roofline("learned", ml, b_iter, lds_iter, n_cu, copy_gbs)
        # Paired error rates: at the workload's priors, and at mean prior 0.003 where failures are seen.
        p3 = np.minimum(p * (0.003 / float(p.mean())), 0.45).astype(np.float32)
        d3 = RelayBPDecoder(H, p3, data_type="float32", algo="relay", seed=0, device=[0], **RELAY)
        dl3 = RelayBPDecoder(H, p3, data_type="float32", algo="relay", seed=0, device=[0], learned_gamma=net,
                             learned_legs=legs, learned_last_damp=damp, **RELAY)
        ler_w = paired_ler({"plain": dec, "learned": dl}, H, p, args.ler_batches, 4096, 101)
        # the same network with one learned leg fewer (one more uniform exploration leg), on the same shots
        fewer = legs - 1 if legs > 1 and not args.no_fewer_legs else 0
        decs3 = {"plain": d3, "learned": dl3}
        if fewer:
            decs3["learned_fewer_legs"] = RelayBPDecoder(H, p3, data_type="float32", algo="relay", seed=0, device=[0],
                                                         learned_gamma=net, learned_legs=fewer,
                                                         learned_last_damp=damp, **RELAY)
        ler_3 = paired_ler(decs3, H, p3, args.ler_batches, 4096, 2001)
        plain_ci = ler_3["plain"]["ler"]
        matched = (plain_ci[1] <= ler_3["learned"]["ler"][0] <= plain_ci[2]) and ler_3["learned"]["mcnemar_p"] >= 0.05
        result["learned"] = {"shots_per_s": round(ml["rate"], 1),
                             "kernel_ms_per_launch": round(sum(ml["kernel_ms"]) / len(ml["kernel_ms"]), 4),
                             "speedup_vs_plain": round(ml["rate"] / m["rate"], 3),
                             "one_at_a_time": seql, "matches_decode": ml["matches"],
                             "mean_iterations": round(ml["batch_iterations"] / args.shots, 3),
                             "learned_legs": legs, "learned_last_damp": damp,
                             "weights": str(wpath.relative_to(ROOT))}
        if fewer:
            dlf = RelayBPDecoder(H, p, data_type="float32", algo="relay", seed=0, device=[0], learned_gamma=net,
                                 learned_legs=fewer, learned_last_damp=damp, **RELAY)
            mf = engine_measure(dlf, syn, args.steps, args.warmup, args.jobs)
            print(f"[bench] learned, {fewer} legs: {mf['rate']:.0f} shots/s", file=sys.stderr, flush=True)
            result["learned_fewer_legs"] = {
                "learned_legs": fewer, "shots_per_s": round(mf["rate"], 1),
                "speedup_vs_plain": round(mf["rate"] / m["rate"], 3), "matches_decode": mf["matches"],
                "mean_iterations": round(mf["batch_iterations"] / args.shots, 3),
                "ler_mean_prior_0.003": ler_3.pop("learned_fewer_legs"),
                "note": "the shipped network with one more uniform exploration leg (fewer learned legs)"}
        result["ler_paired"] = {"workload_priors": ler_w, "mean_prior_0.003": ler_3,
                                "matched": bool(matched),
                                "criterion": "learned LER inside the plain form's Wilson 95% interval at mean "
                                             "prior 0.003 and McNemar p >= 0.05 on the discordant pairs"}
        if matched and ml["rate"] > m["rate"]:
            result["value"] = round(ml["rate"], 1)
            result["value_form"] = "learned"
            result["ms_per_step"] = round(ml["elapsed"] / args.steps * 1e3, 3)
            result["roofline"] = roof_l
            result["plain"]["roofline"] = roof_plain
        else:
            result["learned"]["roofline"] = roof_l
    if not args.no_bb72:
        result["code_capacity_bb72"] = code_capacity_bb72(args)
    if not args.no_batched:
        result["batched"] = batched_three_codes()
    if not args.no_serving:
        result["serving"] = {f"prior_{sc}": serving_runs(args.serving_seconds, sc) for sc in (0.001, 0.003)}
        result["p99_latency_ms"] = {f"prior_{sc}": {str(r["rate"]): r["p99_ms"] for r in sv["rates"]}
                                    for sc, sv in zip((0.001, 0.003), result["serving"].values())}
    if not args.no_cpu:
        lam = utils.compute_log_prior_ratios(p)
        gtab = dec.gamma_table(args.shots).cpu().numpy()
        n2 = min(args.shots, 1024)
        cb, cref = cpu_baseline(dec.graph, syn_np[:n2], lam, gtab)
        cb["sample"] = f"first {n2} of the {args.shots} shots; " + cb["sample"]
        cb["gpu_matches_on_sample"] = all(np.array_equal(ref[k][:n2].cpu().numpy(), cref[k])
                                          for k in ("errors", "iterations", "found_count", "weights"))
        result["cpu_baseline"] = cb
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
