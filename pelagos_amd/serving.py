"""Continuous-batching decode server on one GPU.

Requests (syndrome batches of one of several codes) arrive over time.  The
server keeps every code's decoder, graph layout and launch plans resident on
the device, and a small pool of *slots*: a stream, a workspace and output
buffers sized for the largest batch.  Whenever a slot is free and the HBM
budget allows, all requests waiting for one code (oldest code first) are
merged into one persistent launch on that slot's stream.  Launches on
different slots overlap on the device: a launch's persistent workgroups free
their CUs as its shots finish, and the next launch's workgroups take them.
Each finished launch's per-request results are copied out on the slot's
stream and the slot is recycled.  Latency is measured per request from its
arrival to the moment the host sees its results complete.

Each shot's result does not depend on the other shots of its launch (the
decoder's results are batch-independent), so a request decodes the same alone
and in the mix; tests/test_serving.py checks that bit for bit.

No counterpart in the reference beyond its persistent worker pool and ordered
reassembly (process.py:114-254) and its per-batch-size state cache
(decoder.py:1479-1484).
"""

from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np
import torch

from .decoder import RelayBPDecoder


@dataclass
class Request:
    rid: int
    code: str
    syndromes: torch.Tensor          # [b, C] uint8 on the device
    arrival: float                   # seconds on the server clock
    done: float = -1.0
    errors: Optional[torch.Tensor] = None
    weights: Optional[torch.Tensor] = None
    iterations: Optional[torch.Tensor] = None
    found_count: Optional[torch.Tensor] = None


@dataclass
class _Slot:
    stream: torch.cuda.Stream
    event: torch.cuda.Event
    workspace: torch.Tensor
    errors: torch.Tensor
    weights: torch.Tensor
    iterations: torch.Tensor
    found: torch.Tensor
    busy: bool = False
    reqs: List[Request] = field(default_factory=list)
    footprint: int = 0


class ServerBusy(RuntimeError):
    """The server's queue holds its maximum number of shots: the request was
    not admitted (the caller may retry later)."""


class DecodeServer:
    """decoders: code name -> RelayBPDecoder (one device).  max_batch: shots per
    launch.  slots: concurrent launches (at most the device's hardware queues
    are useful).  hbm_budget: bytes of device memory in-flight work may hold
    (workspace, outputs and per-request result copies).  max_queued_shots:
    shots the queue may hold while they wait for a launch; a request that
    would exceed it is refused with ServerBusy, so a saturated server answers
    at once instead of letting its queue (and every later request's latency)
    grow without bound."""

    def __init__(self, decoders: Dict[str, RelayBPDecoder], max_batch: int = 2048, slots: int = 4,
                 hbm_budget: int = 64 << 30, max_queued_shots: int = 16384):
        if not decoders:
            raise ValueError("no decoders")
        self.decoders = decoders
        self.device = next(iter(decoders.values())).device
        self.max_batch = int(max_batch)
        self.hbm_budget = int(hbm_budget)
        self.max_queued_shots = int(max_queued_shots)
        if self.max_queued_shots < self.max_batch:
            raise ValueError("max_queued_shots must hold at least one full batch")
        self.queued_shots = 0
        vmax = max(d.V for d in decoders.values())
        ws_need = {}
        for name, d in decoders.items():
            # plans for every batch size a launch may have are made lazily; size the
            # workspace for the largest
            ws_need[name] = d.prepare(self.max_batch)
        ws_bytes = max(ws_need.values())
        self.slots: List[_Slot] = []
        for _ in range(int(slots)):
            dev = self.device
            self.slots.append(_Slot(
                stream=torch.cuda.Stream(device=dev),
                event=torch.cuda.Event(),
                workspace=torch.empty(ws_bytes, dtype=torch.uint8, device=dev),
                errors=torch.empty(self.max_batch * vmax, dtype=torch.uint8, device=dev),
                weights=torch.empty(self.max_batch, dtype=torch.float32, device=dev),
                iterations=torch.empty(self.max_batch, dtype=torch.int32, device=dev),
                found=torch.empty(self.max_batch, dtype=torch.int32, device=dev),
            ))
        self.slot_bytes = ws_bytes + self.max_batch * (vmax + 12)
        self.queues: Dict[str, collections.deque] = {name: collections.deque() for name in decoders}
        self.in_flight_bytes = 0
        self.t0 = time.perf_counter()
        self.finished: List[Request] = []
        self._next_id = 0

    # -- client side -------------------------------------------------------------------------------
    def now(self) -> float:
        return time.perf_counter() - self.t0

    def submit(self, code: str, syndromes: torch.Tensor, arrival: Optional[float] = None) -> Request:
        if code not in self.decoders:
            raise KeyError(code)
        dec = self.decoders[code]
        if syndromes.dim() != 2 or syndromes.shape[1] != dec.C:
            raise ValueError(f"{code}: syndromes must be [b, {dec.C}]")
        if syndromes.shape[0] > self.max_batch:
            raise ValueError(f"request of {syndromes.shape[0]} shots exceeds max_batch={self.max_batch}")
        b = int(syndromes.shape[0])
        if self.queued_shots + b > self.max_queued_shots:
            raise ServerBusy(f"{self.queued_shots} shots queued; a request of {b} exceeds "
                             f"max_queued_shots={self.max_queued_shots}")
        r = Request(self._next_id, code, syndromes.to(self.device, torch.uint8).contiguous(),
                    self.now() if arrival is None else arrival)
        self._next_id += 1
        self.queues[code].append(r)
        self.queued_shots += b
        return r

    # -- server side -------------------------------------------------------------------------------
    def _launch(self, slot: _Slot, code: str) -> None:
        dec = self.decoders[code]
        q = self.queues[code]
        reqs, shots = [], 0
        while q and shots + q[0].syndromes.shape[0] <= self.max_batch:
            r = q.popleft()
            reqs.append(r)
            shots += r.syndromes.shape[0]
        self.queued_shots -= shots
        # footprint: workspace and buffers of the slot plus the result copies
        foot = self.slot_bytes + shots * (dec.V + 12)
        slot.busy, slot.reqs, slot.footprint = True, reqs, foot
        self.in_flight_bytes += foot
        # the slot's stream must see the requests' syndromes (written on the submitting stream)
        slot.stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(slot.stream):
            syn = reqs[0].syndromes if len(reqs) == 1 else torch.cat([r.syndromes for r in reqs], 0)
            B = syn.shape[0]
            err = slot.errors[:B * dec.V].view(B, dec.V)  # contiguous [B, V]
            w, it, fc = slot.weights[:B], slot.iterations[:B], slot.found[:B]
            dec.decode_async(syn, err, w, it, fc, slot.workspace, slot.stream)
            o = 0
            for r in reqs:
                b = r.syndromes.shape[0]
                r.errors = err[o:o + b].clone()
                r.weights = w[o:o + b].clone()
                r.iterations = it[o:o + b].clone()
                r.found_count = fc[o:o + b].clone()
                o += b
            slot.event.record(slot.stream)

    def poll(self) -> int:
        """Retire finished launches and start new ones; returns requests finished."""
        n = 0
        t = self.now()
        for s in self.slots:
            if s.busy and s.event.query():
                for r in s.reqs:
                    r.done = t
                    self.finished.append(r)
                n += len(s.reqs)
                self.in_flight_bytes -= s.footprint
                s.busy, s.reqs = False, []
        for s in self.slots:
            if s.busy:
                continue
            waiting = [(q[0].arrival, name) for name, q in self.queues.items() if q]
            if not waiting:
                break
            code = min(waiting)[1]
            dec = self.decoders[code]
            shots = min(self.max_batch, sum(r.syndromes.shape[0] for r in self.queues[code]))
            if self.in_flight_bytes + self.slot_bytes + shots * (dec.V + 12) > self.hbm_budget:
                break
            self._launch(s, code)
        return n

    def pending(self) -> int:
        return sum(len(q) for q in self.queues.values()) + sum(len(s.reqs) for s in self.slots if s.busy)

    def drain(self) -> None:
        while self.pending():
            self.poll()


def poisson_trace(rate_shots_per_s: float, duration_s: float, codes: List[str], seed: int,
                  max_shots: int = 256) -> List[tuple]:
    """Arrival times of requests whose sizes are uniform in 1..max_shots and whose
    code is uniform over ``codes``, at a mean of rate_shots_per_s syndromes/s."""
    rng = np.random.default_rng(seed)
    req_rate = rate_shots_per_s / ((1 + max_shots) / 2)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / req_rate)
        if t >= duration_s:
            return out
        out.append((t, codes[rng.integers(len(codes))], int(rng.integers(1, max_shots + 1))))


def run_trace(server: DecodeServer, trace: List[tuple], syndrome_pool: Dict[str, torch.Tensor],
              policy: str = "reject") -> dict:
    """Replay a trace in real time: submit each request when its arrival time
    comes (its syndromes are slices of a device-resident pool), keep polling,
    and report latency percentiles and throughput.  A request the server
    refuses (ServerBusy) is dropped and counted under ``policy="reject"``, or
    kept by the client and submitted again at each poll under
    ``policy="defer"`` (its latency still counts from its arrival)."""
    if policy not in ("reject", "defer"):
        raise ValueError("policy must be 'reject' or 'defer'")
    server.finished = []
    server.t0 = time.perf_counter()
    offs = {k: 0 for k in syndrome_pool}
    backlog: collections.deque = collections.deque()
    rejected = rejected_shots = deferred = 0
    i = 0
    while i < len(trace) or backlog or server.pending():
        now = server.now()
        while backlog:
            t, code, syn = backlog[0]
            try:
                server.submit(code, syn, arrival=t)
            except ServerBusy:
                break
            backlog.popleft()
        while i < len(trace) and trace[i][0] <= now:
            t, code, b = trace[i]
            pool = syndrome_pool[code]
            o = offs[code] % max(1, pool.shape[0] - b)
            offs[code] += b
            i += 1
            if backlog:
                backlog.append((t, code, pool[o:o + b]))
                deferred += 1
                continue
            try:
                server.submit(code, pool[o:o + b], arrival=t)
            except ServerBusy:
                if policy == "reject":
                    rejected += 1
                    rejected_shots += b
                else:
                    backlog.append((t, code, pool[o:o + b]))
                    deferred += 1
        server.poll()
    end = server.now()
    lat = np.array([r.done - r.arrival for r in server.finished] or [0.0]) * 1e3
    shots = sum(r.syndromes.shape[0] for r in server.finished)
    dur = max(end, trace[-1][0] if trace else end)
    return dict(requests=len(server.finished), shots=int(shots), seconds=round(dur, 3),
                shots_per_s=round(shots / dur, 1), p50_ms=round(float(np.percentile(lat, 50)), 3),
                p99_ms=round(float(np.percentile(lat, 99)), 3), max_ms=round(float(lat.max()), 3),
                policy=policy, rejected_requests=rejected, rejected_shots=int(rejected_shots),
                deferred_requests=deferred)
