"""Continuous-batching decode server on one GPU, with shot-level admission.

Requests (syndrome batches of one of several codes) arrive over time.  The
server keeps every code's decoder, graph layout, gamma rows and a serving
engine (``pel_plan_serve`` / ``pel_serve_launch``) resident on the device: one
persistent launch whose workgroups take units of work from a queue in host
memory while it runs.  ``submit`` turns a request into units at once (one shot
each for a code decoded a shot per workgroup; up to 64 shots for a small code
decoded a shot per wave) and publishes them, so a request's shots start on
whichever workgroups free up next, alongside the shots of earlier requests
still running; a slow shot holds one workgroup, not the requests behind it.
The engine writes each shot's results straight into the request's own output
tensors and raises a done flag per unit in host memory; ``poll`` marks a
request finished when all of its units are done.  The engine's workgroups
leave after ``idle_ms`` without work, and the server launches the engine again
when work arrives and the last launch has ended.

Backpressure: at most ``max_queued_shots`` shots may be submitted and not yet
finished; a request that would exceed that is refused with ServerBusy at once,
so a saturated server answers immediately instead of letting its queue (and
every later request's latency) grow without bound.

HBM budget: with ``hbm_budget_bytes`` the server sizes its device memory to the
budget.  A shot's message state lives in the LDS of the CU decoding it, so the
device memory a server holds is its fixed part (the engine's staging rows,
workspaces and job table) plus each code's result arena (error row and three
counters per shot).  The arenas get the rows the budget leaves after the fixed
part, and the admission bound follows from them (a quarter of the rows, as by
default), so the server admits work while its results fit the budget.

Each shot's result does not depend on the other shots decoded with it (the
decoder's results are batch-independent), so a request decodes the same alone
and in the mix; tests/test_serving.py checks that bit for bit.

No counterpart in the reference beyond its persistent worker pool and ordered
reassembly (process.py:114-254) and its per-batch-size state cache
(decoder.py:1479-1484).
"""

from __future__ import annotations

import collections
import ctypes
import time
from dataclasses import dataclass
from typing import Dict, List, Optional

import numpy as np
import torch

from . import _native
from .decoder import RelayBPDecoder

# pel_serve_unit, as numpy builds a request's units
_UNIT = np.dtype([("job", "<i4"), ("count", "<i4"), ("syn_row_stride", "<i8"), ("syndromes", "<u8"),
                  ("errors", "<u8"), ("weights", "<u8"), ("iterations", "<u8"), ("found_count", "<u8"),
                  ("syn_col_stride", "<i8")])
assert _UNIT.itemsize == ctypes.sizeof(_native.PelServeUnit) == 64


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
    first_unit: int = 0
    n_units: int = 0


class ServerBusy(RuntimeError):
    """The server holds its maximum number of unfinished shots (or a code's
    result arena has no room next to an unfinished request): the request was
    not admitted (the caller may retry later)."""


class _Arena:
    """One code's result buffers, allocated once and handed out as contiguous
    row ranges in ring order: a range is taken at the head, or at row 0 when it
    does not fit before the end, and never over a request still in flight (the
    ranges from the oldest unfinished one to the head).  A finished request's
    rows are reused once the ring comes round to them again."""

    def __init__(self, rows: int, V: int, device):
        self.rows = rows
        self.errors = torch.empty(rows, V, dtype=torch.uint8, device=device)
        self.weights = torch.empty(rows, dtype=torch.float32, device=device)
        self.iterations = torch.empty(rows, dtype=torch.int32, device=device)
        self.found_count = torch.empty(rows, dtype=torch.int32, device=device)
        self.head = 0
        self.live: collections.deque = collections.deque()  # (start, request) in allocation order

    def take(self, b: int, r: "Request") -> Optional[int]:
        while self.live and self.live[0][1].done >= 0:
            self.live.popleft()
        n, h = self.rows, self.head
        if not self.live:
            start = h if h + b <= n else 0
        else:
            s0 = self.live[0][0]
            if h > s0:    # in flight: [s0, h); free: [h, n) and [0, s0)
                start = h if h + b <= n else (0 if b <= s0 else None)
            elif h < s0:  # in flight: [s0, n) and [0, h); free: [h, s0)
                start = h if h + b <= s0 else None
            else:         # the ring is full
                start = None
        if start is None or b > n:
            return None
        self._undo = self.head
        self.head = start + b
        self.live.append((start, r))
        return start

    def undo(self) -> None:
        """Return the range the last take handed out (its request was not admitted)."""
        self.live.pop()
        self.head = self._undo


class DecodeServer:
    """decoders: code name -> RelayBPDecoder (one device, broadcast gamma draws,
    shared priors).  max_batch: largest request (shots).  max_queued_shots:
    shots submitted and not yet finished.  idle_ms: how long the engine's
    workgroups wait for work before the launch ends.  queue_units: entries of
    the unit ring; an entry is reused once its unit is done, so a slow shot
    refuses new work only after queue_units later units (default 2^18: seconds
    of work at full rate).

    Syndromes are read in place, with their strides: device-resident uint8
    (or bool) tensors, complete (whatever produced them finished) when
    submitted; anything else is converted and the current stream synchronised
    first.  Nothing is allocated per request: a request's
    result tensors are views of its code's result arena (``result_shots`` rows,
    default 4 x max_queued_shots), complete once ``poll`` has marked the request
    done and valid until about ``result_shots`` more shots of that code have
    been submitted (clone them to keep them longer).

    hbm_budget_bytes: size the device memory to this budget (see the module
    docstring); result_shots and max_queued_shots, unless given, then follow
    from it (max_queued_shots at most queue_units), and a budget below the
    fixed part or below one full request is refused.  ``device_bytes`` reports
    what the server holds."""

    def __init__(self, decoders: Dict[str, RelayBPDecoder], max_batch: int = 2048,
                 max_queued_shots: Optional[int] = None, idle_ms: float = 2.0, result_shots: Optional[int] = None,
                 queue_units: int = 1 << 18, hbm_budget_bytes: Optional[int] = None):
        if not decoders:
            raise ValueError("no decoders")
        if len(decoders) > _native.PEL_MAX_JOBS:
            raise ValueError(f"at most {_native.PEL_MAX_JOBS} codes")
        self.decoders = decoders
        self.names = list(decoders)
        self.device = next(iter(decoders.values())).device
        if any(d.device != self.device for d in decoders.values()):
            raise ValueError("all decoders must be on one device")
        for name, d in decoders.items():
            if d.gamma_table(1).dim() != 2:
                raise ValueError(f"{name}: serving needs broadcast gamma draws (gamma_draw != 'per_row')")
        self.max_batch = int(max_batch)
        if self.max_batch < 1:
            raise ValueError("max_batch must be >= 1")
        budget = None if hbm_budget_bytes is None else int(hbm_budget_bytes)
        if max_queued_shots is None and budget is None:
            max_queued_shots = 16384
        if max_queued_shots is not None:
            self._set_admission(int(max_queued_shots), queue_units)
        self.idle_us = int(round(float(idle_ms) * 1000))
        n = len(self.names)
        L = _native.lib()
        self._graphs = (ctypes.c_void_p * n)(*[decoders[k]._handle.h.value for k in self.names])
        flags = (ctypes.c_int32 * n)(*[
            _native.PEL_PLAN_LEARNED if (decoders[k].learned_params is not None and decoders[k].learned_legs > 0)
            else 0 for k in self.names])
        self.plan = _native.PelJobsPlan()
        dev = self.device
        self._queue = ctypes.c_void_p()
        with torch.cuda.device(dev):
            _native.check(L.pel_plan_serve(self._graphs, flags, n, ctypes.byref(self.plan)), "pel_plan_serve")
            self._staging, self._ws, params, batches = {}, {}, [], []
            self.job = {}
            for i, name in enumerate(self.names):
                d = decoders[name]
                rows = int(self.plan.grid) * (16 if int(self.plan.group[i]) == 64 else 1)
                d.prepare(rows)
                st = dict(syndromes=torch.zeros(rows, d.C, dtype=torch.uint8, device=dev),
                          errors=torch.empty(rows, d.V, dtype=torch.uint8, device=dev),
                          weights=torch.empty(rows, dtype=torch.float32, device=dev),
                          iterations=torch.empty(rows, dtype=torch.int32, device=dev),
                          found_count=torch.empty(rows, dtype=torch.int32, device=dev))
                self._staging[name] = st
                self._ws[name] = torch.empty(max(int(self.plan.workspace_bytes[i]), 1), dtype=torch.uint8, device=dev)
                prm, bat = d.job_args(st["syndromes"], st["errors"], st["weights"], st["iterations"],
                                      st["found_count"], self._ws[name])
                bat.shot_clock = None
                params.append(prm)
                batches.append(bat)
                self.job[name] = (i, 1 if int(self.plan.group[i]) != 64 else _native.PEL_SERVE_WAVE_SHOTS)
            self._params = (_native.PelParams * n)(*params)
            self._batches = (_native.PelBatch * n)(*batches)
            self._table = torch.empty(int(self.plan.table_bytes), dtype=torch.uint8, device=dev)
            fixed = sum(t.numel() * t.element_size() for st in self._staging.values() for t in st.values())
            fixed += sum(t.numel() for t in self._ws.values()) + self._table.numel()
            row_bytes = sum(decoders[k].V + 12 for k in self.names)  # error row + weight, iterations, found
            if budget is not None:
                room = (budget - fixed) // row_bytes
                if result_shots is None:
                    result_shots = room
                if result_shots < self.max_batch or result_shots > room:
                    raise ValueError(f"hbm_budget_bytes={budget}: the fixed part takes {fixed} B and each result "
                                     f"row {row_bytes} B, room for {max(room, 0)} rows; a full request needs "
                                     f"{self.max_batch}")
                if max_queued_shots is None:
                    self._set_admission(max(self.max_batch, min(int(result_shots) // 4, int(queue_units))),
                                        queue_units)
            rs = int(result_shots) if result_shots is not None else 4 * self.max_queued_shots
            if rs < self.max_batch:
                raise ValueError("result_shots must hold at least one full request")
            _native.check(L.pel_serve_queue_create(int(queue_units), ctypes.byref(self._queue)),
                          "pel_serve_queue_create")
            cap = ctypes.c_int32()
            _native.check(L.pel_serve_queue_info(self._queue, None, ctypes.byref(cap), None),
                          "pel_serve_queue_info")
            self._cap = int(cap.value)
            self._lib = L
            self._arena = {k: _Arena(rs, decoders[k].V, dev) for k in self.names}
            self.result_shots = rs
            self._device_bytes = fixed + rs * row_bytes
            self.stream = torch.cuda.Stream(device=dev)
        self._ended: Optional[torch.cuda.Event] = None
        self.launches = 0
        self.queued_shots = 0
        self.outstanding: collections.deque = collections.deque()
        self.t0 = time.perf_counter()
        self.finished: List[Request] = []
        self._next_id = 0

    def _set_admission(self, max_queued_shots: int, queue_units: int) -> None:
        self.max_queued_shots = max_queued_shots
        if self.max_queued_shots < self.max_batch:
            raise ValueError("max_queued_shots must hold at least one full request of max_batch >= 1 shots")
        if queue_units < self.max_queued_shots:
            raise ValueError("queue_units must be at least max_queued_shots")

    def device_bytes(self) -> int:
        """Device memory the server holds: staging rows, workspaces, the job
        table and the result arenas (the engine's unit queue is in host memory)."""
        return self._device_bytes

    def roles(self) -> Dict[str, dict]:
        """Each code's role in the engine: threads per shot, placement, learned."""
        p = self.plan
        return {n: dict(group=int(p.group[i]), placement=int(p.placement[i]), learned=bool(p.learned[i]))
                for i, n in enumerate(self.names)}

    # -- client side -------------------------------------------------------------------------------
    def now(self) -> float:
        return time.perf_counter() - self.t0

    def submit(self, code: str, syndromes: torch.Tensor, arrival: Optional[float] = None) -> Request:
        if self._queue is None:
            raise RuntimeError("the server is closed")
        if code not in self.decoders:
            raise KeyError(code)
        dec = self.decoders[code]
        if syndromes.dim() != 2 or syndromes.shape[1] != dec.C:
            raise ValueError(f"{code}: syndromes must be [b, {dec.C}]")
        b = int(syndromes.shape[0])
        if b > self.max_batch:
            raise ValueError(f"request of {b} shots exceeds max_batch={self.max_batch}")
        if self.queued_shots + b > self.max_queued_shots:
            raise ServerBusy(f"{self.queued_shots} shots unfinished; a request of {b} exceeds "
                             f"max_queued_shots={self.max_queued_shots}")
        if syndromes.device == self.device and syndromes.dtype in (torch.uint8, torch.bool):
            # read in place, strided as it is: no work is put on a stream (a copy kernel would wait
            # for the engine's workgroups to leave)
            syn = syndromes.view(torch.uint8) if syndromes.dtype == torch.bool else syndromes
        else:
            syn = syndromes.to(self.device, torch.uint8)
            torch.cuda.current_stream(self.device).synchronize()  # the converted copy is complete
        r = Request(self._next_id, code, syn, self.now() if arrival is None else arrival)
        ar = self._arena[code]
        o = ar.take(b, r) if b > 0 else 0
        if o is None:
            raise ServerBusy(f"{code}: no room in the result arena next to unfinished requests")
        r.errors, r.weights = ar.errors[o:o + b], ar.weights[o:o + b]
        r.iterations, r.found_count = ar.iterations[o:o + b], ar.found_count[o:o + b]
        self._next_id += 1
        if b == 0:
            r.done = r.arrival
            self.finished.append(r)
            return r
        job, per_unit = self.job[code]
        starts = np.arange(0, b, per_unit, dtype=np.int64)
        u = np.zeros(starts.size, dtype=_UNIT)
        u["job"] = job
        u["count"] = np.minimum(per_unit, b - starts)
        rs, cs = syn.stride()
        u["syn_row_stride"] = rs
        u["syn_col_stride"] = cs
        u["syndromes"] = syn.data_ptr() + starts * rs
        u["errors"] = r.errors.data_ptr() + starts * dec.V
        u["weights"] = r.weights.data_ptr() + 4 * starts
        u["iterations"] = r.iterations.data_ptr() + 4 * starts
        u["found_count"] = r.found_count.data_ptr() + 4 * starts
        first = ctypes.c_int64()
        rc = _native.lib().pel_serve_publish(self._queue, u.ctypes.data_as(ctypes.POINTER(_native.PelServeUnit)),
                                             int(starts.size), ctypes.byref(first))
        if rc != _native.PEL_OK:
            ar.undo()
            if rc == _native.PEL_EBUSY:
                raise ServerBusy(_native.lib().pel_last_error().decode(errors="replace"))
            _native.check(rc, "pel_serve_publish")
        r.first_unit, r.n_units = int(first.value), int(starts.size)
        self.outstanding.append(r)
        self.queued_shots += b
        self._ensure_running()
        return r

    # -- server side -------------------------------------------------------------------------------
    def _ensure_running(self) -> None:
        """Launch the engine if no launch is running (the last one may have left
        while units were published)."""
        if self._ended is not None and not self._ended.query():
            return
        with torch.cuda.device(self.device):
            _native.check(_native.lib().pel_serve_launch(
                self._graphs, self._params, self._batches, len(self.names), ctypes.byref(self.plan),
                ctypes.c_void_p(self._table.data_ptr()), self._queue, self.idle_us,
                ctypes.c_void_p(self.stream.cuda_stream)), "pel_serve_launch")
            ev = torch.cuda.Event()
            ev.record(self.stream)
        self._ended = ev
        self.launches += 1

    def _is_done(self, r: Request) -> bool:
        return int(self._lib.pel_serve_done_count(self._queue, r.first_unit, r.n_units)) == r.n_units

    def poll(self) -> int:
        """Mark finished requests (and relaunch the engine if it left with work
        outstanding); returns the number of requests finished."""
        if not self.outstanding:
            return 0
        t = self.now()
        keep: collections.deque = collections.deque()
        n = 0
        for r in self.outstanding:
            if self._is_done(r):
                r.done = t
                self.finished.append(r)
                self.queued_shots -= int(r.syndromes.shape[0])
                n += 1
            else:
                keep.append(r)
        self.outstanding = keep
        if self.outstanding:
            self._ensure_running()
        return n

    def pending(self) -> int:
        return len(self.outstanding)

    def drain(self, timeout_s: float = 600.0) -> None:
        t_end = time.perf_counter() + timeout_s
        while self.pending():
            self.poll()
            if time.perf_counter() > t_end:
                raise TimeoutError(f"{self.pending()} requests still unfinished after {timeout_s} s")

    def drain_requests(self, reqs: List[Request], timeout_s: float = 600.0) -> None:
        """Poll until every request in ``reqs`` is finished."""
        t_end = time.perf_counter() + timeout_s
        while any(r.done < 0 for r in reqs):
            self.poll()
            if time.perf_counter() > t_end:
                raise TimeoutError(f"requests still unfinished after {timeout_s} s")

    def quiesce(self) -> None:
        """End the running launch (if any) and wait for it: its workgroups
        finish every unit published so far, then leave instead of waiting for
        more.  The next submit launches the engine again."""
        if self._ended is None:
            return
        L = _native.lib()
        _native.check(L.pel_serve_set_stop(self._queue, 1), "pel_serve_set_stop")
        try:
            self._ended.synchronize()
        finally:
            _native.check(L.pel_serve_set_stop(self._queue, 0), "pel_serve_set_stop")

    def close(self) -> None:
        if self._queue is None:
            return
        self.quiesce()
        _native.check(_native.lib().pel_serve_queue_destroy(self._queue), "pel_serve_queue_destroy")
        self._queue = None

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self) -> "DecodeServer":
        return self

    def __exit__(self, *exc) -> None:
        self.close()


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
