"""Several batches -- of one code or several -- in one persistent launch, and
a launch-shape autotuner.

``JobLaunch`` decodes up to PEL_MAX_JOBS jobs (a job: one decoder and a batch
of its syndromes) with a single ``pel_decode_jobs`` launch.  The library plans
each job's role: a small code (the [[72,12,6]] code) runs a shot per wave,
sixteen per workgroup; a large code runs a shot per workgroup, its state in
LDS when it fits and spilled to the workspace when it does not.  The launch's
workgroups take the jobs in order and move to the next job as soon as the
current one has no shot left, so one job's slow last shots overlap the next
jobs' work.  Buffers (syndromes in, results out, workspaces) are allocated
once; callers refill ``syndromes[name]`` in place between launches.  The
launch can be captured in a HIP graph: its work is kernels only (the queue
counters are cleared by a kernel, not a memset).

``MultiCodeBatch`` is the same over a dict of codes (one job per code).
"""

from __future__ import annotations

import ctypes
from typing import Dict, Optional, Sequence, Tuple

import torch

from . import _native
from .decoder import RelayBPDecoder


class JobLaunch:
    """jobs: a sequence of (name, decoder, shots); names must be distinct.  All
    decoders must be on one device, with broadcast gamma draws."""

    def __init__(self, jobs: Sequence[Tuple[str, RelayBPDecoder, int]]):
        if not 1 <= len(jobs) <= _native.PEL_MAX_JOBS:
            raise ValueError(f"between 1 and {_native.PEL_MAX_JOBS} jobs")
        names = [n for n, _, _ in jobs]
        if len(set(names)) != len(names):
            raise ValueError("job names must be distinct")
        self.names = names
        self.decoders = {n: d for n, d, _ in jobs}
        self.shots = {n: int(b) for n, _, b in jobs}
        dev = jobs[0][1].device
        if any(d.device != dev for _, d, _ in jobs):
            raise ValueError("all jobs must be on one device")
        self.device = dev
        n = len(jobs)
        L = _native.lib()
        self._graphs = (ctypes.c_void_p * n)(*[d._handle.h.value for _, d, _ in jobs])
        shots = (ctypes.c_int32 * n)(*[int(b) for _, _, b in jobs])
        flags = (ctypes.c_int32 * n)(*[
            _native.PEL_PLAN_LEARNED if (d.learned_params is not None and d.learned_legs > 0) else 0
            for _, d, _ in jobs])
        self.plan = _native.PelJobsPlan()
        with torch.cuda.device(dev):
            _native.check(L.pel_plan_jobs(self._graphs, shots, flags, n, ctypes.byref(self.plan)), "pel_plan_jobs")
        self.syndromes: Dict[str, torch.Tensor] = {}
        self.out: Dict[str, Dict[str, torch.Tensor]] = {}
        self._ws: Dict[str, torch.Tensor] = {}
        params, batches = [], []
        for i, (name, dec, B) in enumerate(jobs):
            dec.prepare(B)
            self._ws[name] = torch.empty(max(int(self.plan.workspace_bytes[i]), 1), dtype=torch.uint8, device=dev)
            self.syndromes[name] = torch.zeros(B, dec.C, dtype=torch.uint8, device=dev)
            o = dict(errors=torch.empty(B, dec.V, dtype=torch.uint8, device=dev),
                     weights=torch.empty(B, dtype=torch.float32, device=dev),
                     iterations=torch.empty(B, dtype=torch.int32, device=dev),
                     found_count=torch.empty(B, dtype=torch.int32, device=dev))
            self.out[name] = o
            prm, bat = dec.job_args(self.syndromes[name], o["errors"], o["weights"], o["iterations"],
                                    o["found_count"], self._ws[name])
            params.append(prm)
            batches.append(bat)
        self._params = (_native.PelParams * n)(*params)
        self._batches = (_native.PelBatch * n)(*batches)
        self._table = torch.empty(int(self.plan.table_bytes), dtype=torch.uint8, device=dev)
        self.graph: Optional[torch.cuda.CUDAGraph] = None
        self.kernel_events: Optional[list] = None

    def roles(self) -> Dict[str, dict]:
        """Each job's planned role: threads per shot, placement, workspace."""
        p = self.plan
        return {n: dict(group=int(p.group[i]), placement=int(p.placement[i]), learned=bool(p.learned[i]),
                        workspace_bytes=int(p.workspace_bytes[i])) for i, n in enumerate(self.names)}

    def launch(self, stream: Optional[torch.cuda.Stream] = None) -> None:
        """One launch over every job, ordered on ``stream`` (default: the
        current stream).  With ``kernel_events`` a list, a (start, end) pair of
        HIP events recorded around the decode kernel is appended to it."""
        cur = stream if stream is not None else torch.cuda.current_stream(self.device)
        ev = self.kernel_events
        b0 = self._batches[0]
        if ev is not None:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(cur)
            e1.record(cur)
            b0.ev_begin, b0.ev_end = e0.cuda_event, e1.cuda_event
        try:
            with torch.cuda.device(self.device):
                _native.check(_native.lib().pel_decode_jobs(self._graphs, self._params, self._batches,
                                                            len(self.names), ctypes.byref(self.plan),
                                                            ctypes.c_void_p(self._table.data_ptr()),
                                                            ctypes.c_void_p(cur.cuda_stream)), "pel_decode_jobs")
        finally:
            b0.ev_begin = b0.ev_end = None
        if ev is not None:
            ev.append((e0, e1))

    def run_eager(self) -> None:
        self.launch()

    def capture(self) -> None:
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        cap = torch.cuda.Stream(device=self.device)
        with torch.cuda.graph(g, stream=cap):
            self.launch(cap)
        torch.cuda.synchronize(self.device)
        self.graph = g

    def replay(self) -> None:
        if self.graph is None:
            self.capture()
        self.graph.replay()


class MultiCodeBatch(JobLaunch):
    """One job per code: decoders and shots are dicts keyed by code name."""

    def __init__(self, decoders: Dict[str, RelayBPDecoder], shots: Dict[str, int]):
        super().__init__([(name, dec, shots[name]) for name, dec in decoders.items()])
