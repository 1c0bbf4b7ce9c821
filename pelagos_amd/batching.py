"""Several codes in one graph launch, and a launch-shape autotuner.

``MultiCodeBatch`` captures one persistent decode per code (each on its own
stream, forked from and joined back to the capture stream) into a single HIP
graph.  One ``replay()`` then decodes every code's batch: the per-code
launches run side by side on the device, each retiring converged shots from
its own work queue, and the host issues one graph launch instead of three
decode calls.  Buffers (syndromes in, results out, workspaces) are allocated
once at capture; callers refill ``syndromes[name]`` in place between replays.
The captured work contains kernels only (the decode library clears its queue
counters with a kernel, not a memset).

``autotune`` times the launch shapes ``pel_plan_decode`` allows for one code
and batch size (workgroup size, re-queue threshold) on the caller's
syndromes and keeps the fastest in a JSON cache keyed by the graph's sparsity
hash, the batch size and the device name, like the reference's per-problem
autotune cache (autotune.py:130-133, 240-259).
"""

from __future__ import annotations

import json
import os
from pathlib import Path
from typing import Dict, Iterable, Optional, Tuple

import torch

from .decoder import RelayBPDecoder


class MultiCodeBatch:
    def __init__(self, decoders: Dict[str, RelayBPDecoder], shots: Dict[str, int]):
        self.decoders = decoders
        self.shots = dict(shots)
        dev = next(iter(decoders.values())).device
        self.device = dev
        self.syndromes: Dict[str, torch.Tensor] = {}
        self.out: Dict[str, Dict[str, torch.Tensor]] = {}
        self._ws: Dict[str, torch.Tensor] = {}
        for name, dec in decoders.items():
            B = self.shots[name]
            ws = dec.prepare(B)
            self._ws[name] = torch.empty(ws, dtype=torch.uint8, device=dev)
            self.syndromes[name] = torch.zeros(B, dec.C, dtype=torch.uint8, device=dev)
            self.out[name] = dict(
                errors=torch.empty(B, dec.V, dtype=torch.uint8, device=dev),
                weights=torch.empty(B, dtype=torch.float32, device=dev),
                iterations=torch.empty(B, dtype=torch.int32, device=dev),
                found_count=torch.empty(B, dtype=torch.int32, device=dev),
            )
        self._streams = {name: torch.cuda.Stream(device=dev) for name in decoders}
        self.graph: Optional[torch.cuda.CUDAGraph] = None

    def _issue(self, main: torch.cuda.Stream) -> None:
        for name, dec in self.decoders.items():
            s = self._streams[name]
            s.wait_stream(main)
            o = self.out[name]
            dec.decode_async(self.syndromes[name], o["errors"], o["weights"], o["iterations"], o["found_count"],
                             self._ws[name], s)
        for s in self._streams.values():
            main.wait_stream(s)

    def run_eager(self) -> None:
        """The same launches without the graph (on the current stream's timeline)."""
        self._issue(torch.cuda.current_stream(self.device))

    def capture(self) -> None:
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        cap = torch.cuda.Stream(device=self.device)
        with torch.cuda.graph(g, stream=cap):
            self._issue(cap)
        torch.cuda.synchronize(self.device)
        self.graph = g

    def replay(self) -> None:
        if self.graph is None:
            self.capture()
        self.graph.replay()


def _cache_path() -> Path:
    base = os.environ.get("PELAGOS_AMD_TUNE_CACHE")
    if base:
        return Path(base)
    return Path(os.path.expanduser("~")) / ".cache" / "pelagos_amd_tune.json"


def _key(dec: RelayBPDecoder, B: int) -> str:
    name = torch.cuda.get_device_name(dec.device)
    return f"{dec.graph.sparsity_hash}|B={B}|{name}"


def autotune(dec: RelayBPDecoder, syndromes: torch.Tensor, blocks: Iterable[int] = (256, 512, 1024),
             requeue: Iterable[int] = (0, 16, 32), reps: int = 3, cache: Optional[Path] = None,
             force: bool = False) -> Tuple[int, int, float]:
    """Pick (block, requeue_below) for this decoder and batch size by timing
    each allowed shape (kernel time from HIP events, best of ``reps``); apply
    it to ``dec`` and return (block, requeue_below, milliseconds)."""
    B = int(syndromes.shape[0])
    path = cache or _cache_path()
    key = _key(dec, B)
    table = {}
    if path.exists():
        try:
            table = json.loads(path.read_text())
        except (OSError, ValueError):
            table = {}
    if key in table and not force:
        blk, rq, ms = table[key]
        _apply(dec, blk, rq)
        return blk, rq, ms
    best = None
    for blk in blocks:
        for rq in requeue:
            try:
                _apply(dec, blk, rq)
                dec.decode(syndromes)  # plan + warm-up
            except (ValueError, RuntimeError):
                continue  # shape not allowed for this graph (LDS, block size)
            times = []
            for _ in range(reps):
                ev = []
                dec.kernel_events = ev
                dec.decode(syndromes)
                torch.cuda.synchronize(dec.device)
                dec.kernel_events = None
                times.append(ev[0][0].elapsed_time(ev[0][1]))
            ms = min(times)
            if best is None or ms < best[2]:
                best = (blk, rq, ms)
    if best is None:
        raise RuntimeError("no launch shape could run this decode")
    _apply(dec, best[0], best[1])
    table[key] = list(best)
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps(table, indent=1))
    except OSError:
        pass
    return best


def _apply(dec: RelayBPDecoder, block: int, requeue_below: int) -> None:
    dec._block_size = int(block)
    dec.requeue_below = int(requeue_below)
    dec._plans.clear()
