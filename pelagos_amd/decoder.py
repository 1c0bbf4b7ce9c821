"""RelayBPDecoder: the reference's Relay-BP-S decoder surface on MI355X.

Same constructor (src/pelagos/decoder.py:271-307), same ``decode`` contract
(decoder.py:2317-2449) and the same exception types and messages for the
validations the reference performs.  Underneath, one call of ``decode`` is a
single persistent HIP launch (``pel_decode``, include/pelagos_relaybp.h):
every shot runs its whole pre-iter leg and relay legs inside one workgroup,
so there is no per-iteration host synchronisation.

PyTorch supplies device memory, streams and the per-leg memory-strength draws
(the reference's exact ``torch.rand`` calls on the decoding device,
utils.py:128-225); all decoding arithmetic runs in the HIP library.
"""

from __future__ import annotations

import ctypes
import json
import os
import threading
from pathlib import Path
from typing import Dict, List, Optional, Tuple, Union

import numpy as np
import torch
from scipy.sparse import csr_matrix

from . import _native
from .graph import CSRGraph
from .utils import (
    GAMMA_DRAW_MODES,
    UNCONVERGED_CORRECTIONS,
    bitpack_errors,
    compute_log_prior_ratios,
    explicit_leg_rows,
    sample_gamma_per_variable,
    sample_gamma_uniform,
    validate_csr_matrix,
    validate_error_priors,
)

_DTYPE_ALIASES = {
    "float": "float32",
    "half": "float16",
    "double": "float64",
    "fp16": "float16",
    "fp32": "float32",
    "fp64": "float64",
}
_PERF_ALIASES = {
    "default": "default",
    "row": "default",
    "rowwise": "default",
    "throughput": "throughput",
    "btile": "throughput",
}


def alpha_schedule(alpha: Optional[float], scaling: float, n: int) -> np.ndarray:
    """alpha for in-leg iteration t = 0..n-1, as the float32 kernel scalar
    (decoder.py:2649-2660): 1 - 2^{-(t+1)/s} when alpha == 0, 1 when alpha < 0
    or None, else alpha."""
    out = np.ones(max(n, 1), dtype=np.float32)
    if alpha is None:
        return out
    a = float(alpha)
    for t in range(n):
        if a == 0.0:
            s = float(scaling or 1.0)
            out[t] = np.float32(1.0 - (2.0 ** (-((t + 1) / s))))
        elif a < 0.0:
            out[t] = 1.0
        else:
            out[t] = np.float32(a)
    return out


class _GraphHandle:
    """Owns a pel_graph on one device."""

    def __init__(self, graph: CSRGraph, device: torch.device):
        self.device = device
        L = _native.lib()
        h = ctypes.c_void_p()
        with torch.cuda.device(device):
            _native.check(
                L.pel_graph_create(graph.chk_ptr.ctypes.data, graph.edge_var.ctypes.data,
                                   graph.C, graph.V, graph.E, ctypes.byref(h)),
                "pel_graph_create",
            )
        self.h = h

    def plan(self, B: int, block: int = 0, placement: int = -1, learned: bool = False,
             legwise: bool = False) -> _native.PelPlan:
        p = _native.PelPlan()
        flags = (_native.PEL_PLAN_LEARNED if learned else 0) | (_native.PEL_PLAN_LEGWISE if legwise else 0)
        with torch.cuda.device(self.device):
            _native.check(_native.lib().pel_plan_decode(self.h, int(B), int(block), int(placement), flags,
                                                        ctypes.byref(p)),
                          "pel_plan_decode")
        return p

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                _native.lib().pel_graph_destroy(h)
            except Exception:
                pass
            self.h = None


def _ptr(t: Optional[torch.Tensor]) -> Optional[int]:
    return None if t is None else t.data_ptr()


_TUNE_LOCK = threading.Lock()


def tune_cache_path() -> Path:
    """The launch-shape cache: PELAGOS_AMD_TUNE_CACHE, else
    ~/.cache/pelagos_amd_tune.json."""
    base = os.environ.get("PELAGOS_AMD_TUNE_CACHE")
    return Path(base) if base else Path(os.path.expanduser("~")) / ".cache" / "pelagos_amd_tune.json"


def _load_tune_cache() -> dict:
    try:
        return json.loads(tune_cache_path().read_text())
    except (OSError, ValueError):
        return {}


def _save_tune_cache(table: dict) -> None:
    """Write atomically (a temporary file renamed over the cache), so that
    concurrent decoders never read a half-written file."""
    path = tune_cache_path()
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(f".{os.getpid()}.tmp")
        tmp.write_text(json.dumps(table, indent=1))
        os.replace(tmp, path)
    except OSError:
        pass


class RelayBPDecoder:
    """Relay-BP-S min-sum decoder with disordered memory strengths on one MI355X.

    Arguments and defaults follow the reference constructor
    (decoder.py:271-307).  Options the MI355X path does not implement raise
    ``NotImplementedError`` at construction instead of being ignored:
    ``data_type`` other than float32, layered scheduling, more than one GPU.
    ``block_size`` and ``state_placement`` (additions) force the workgroup size
    and the LDS placement bits of the decode kernel (see ``pel_plan_decode``).
    ``learned_gamma`` (addition: a :class:`pelagos_amd.learned.GammaMLP` or its
    parameter vector) with ``learned_legs = L > 0`` gives relay legs 0..L-1
    per-variable memory strengths from the network, evaluated inside the decode
    kernel; the remaining legs use the uniform draws as before.
    ``learned_last_damp`` in [0, 1] scales the network's shift on leg L-1 by
    1 - learned_last_damp (a milder last learned leg).
    ``autotune`` (addition): on the first decode of a batch size the launch
    shapes the library allows (workgroup size, re-queue threshold) are timed on
    the caller's syndromes and the fastest is kept, in memory and in a JSON
    cache keyed by the problem (``tune_cache_path()``), as the reference tunes
    its kernels on first sight of a problem key (decoder.py:905-969,
    autotune.py:240-259).  None: on unless the environment sets
    PELAGOS_AMD_AUTOTUNE=0.  Results never depend on the launch shape.
    """

    def __init__(
        self,
        H_csr: csr_matrix,
        error_priors: np.ndarray,
        pre_iter: int = 80,
        num_sets: int = 100,
        set_max_iter: int = 60,
        gamma0: float = 0.65,
        gamma_dist_interval: Tuple[float, float] = (-0.24, 0.66),
        stop_nconv: int = 5,
        normalized_min_sum_alpha: Optional[float] = None,
        offset_min_sum_beta: Optional[float] = None,
        data_type: str = "float32",
        device: Union[str, List[int], int] = [0],
        seed: int = 1234,
        gamma_draw: str = "broadcast",
        bitpack_output: bool = False,
        algo: Optional[str] = None,
        perf: Optional[str] = None,
        alpha_iteration_scaling_factor: Optional[float] = 1.0,
        explicit_gammas: Optional[np.ndarray] = None,
        gamma_dist_interval_by_error: Optional[np.ndarray] = None,
        fast_max_unconverged_pct: Optional[float] = None,
        scheduling: str = "flood-fill",
        layered_k_iters: int = 0,
        layered_block_policy: str = "contiguous",
        layered_num_blocks: Optional[int] = None,
        layered_target_edges_per_block: Optional[int] = None,
        layered_block_seed: int = 42,
        unconverged_correction: str = "first_leg",
        block_size: Optional[int] = None,
        state_placement: Optional[int] = None,
        learned_gamma=None,
        learned_legs: int = 0,
        learned_last_damp: float = 0.0,
        autotune: Optional[bool] = None,
    ):
        # Validation in the reference's order and wording (decoder.py:362-512).
        if unconverged_correction not in UNCONVERGED_CORRECTIONS:
            raise ValueError(
                f"unconverged_correction must be one of {UNCONVERGED_CORRECTIONS} "
                f"(got {unconverged_correction!r})"
            )
        self.unconverged_correction = unconverged_correction
        validate_csr_matrix(H_csr)
        error_priors = validate_error_priors(error_priors)
        if normalized_min_sum_alpha is not None and offset_min_sum_beta is not None:
            raise ValueError("Cannot specify both normalized_min_sum_alpha and offset_min_sum_beta")
        if normalized_min_sum_alpha is not None and not (0 <= normalized_min_sum_alpha <= 1):
            raise ValueError("normalized_min_sum_alpha must be in [0, 1]")
        if offset_min_sum_beta is not None and offset_min_sum_beta < 0:
            raise ValueError("offset_min_sum_beta must be >= 0")
        self.H_csr = H_csr
        self.check_permutation = None
        self.error_priors = error_priors
        self.pre_iter = int(pre_iter)
        self.num_sets = int(num_sets)
        self.set_max_iter = int(set_max_iter)
        self.gamma0 = gamma0
        self.gamma_dist_interval = gamma_dist_interval
        self.stop_nconv = int(stop_nconv)
        if self.stop_nconv < 1:
            raise ValueError(f"stop_nconv must be >= 1, got {stop_nconv}")
        if min(self.pre_iter, self.num_sets, self.set_max_iter) < 0:
            raise ValueError("pre_iter, num_sets and set_max_iter must be >= 0")
        self.fast_max_unconverged_pct = fast_max_unconverged_pct
        if fast_max_unconverged_pct is not None:
            if not (0.0 <= fast_max_unconverged_pct <= 1.0):
                raise ValueError(
                    "fast_max_unconverged_pct must be between 0.0 and 1.0, got "
                    f"{fast_max_unconverged_pct}"
                )
        self.normalized_min_sum_alpha = normalized_min_sum_alpha
        self.offset_min_sum_beta = offset_min_sum_beta

        data_type = _DTYPE_ALIASES.get(data_type, data_type)
        if data_type not in ("float16", "float32", "float64"):
            raise ValueError(
                f"data_type must be 'float16', 'float32', or 'float64', got {data_type}"
            )
        self.dtype_str = data_type
        self.dtype = {"float16": torch.float16, "float32": torch.float32, "float64": torch.float64}[data_type]
        # float16 / float64 run the per-edge kernel (pel_decode_typed), float32 the fast one
        self._typed = {"float16": _native.PEL_DTYPE_F16, "float64": _native.PEL_DTYPE_F64}.get(data_type)

        self.device_ids = self._parse_device_spec(device)
        self.num_gpus = len(self.device_ids)
        if not torch.cuda.is_available():
            raise RuntimeError(
                "GPU not available (install a GPU-enabled PyTorch build: CUDA or ROCm)"
            )
        if self.num_gpus != 1:
            raise NotImplementedError(
                "one decoder drives one GPU; shard shots across decoders, one per GPU"
            )
        self._backend = "rocm"
        self.device_type = "cuda"
        self.device = torch.device(self.device_type, self.device_ids[0])

        self.seed = seed
        if gamma_draw not in GAMMA_DRAW_MODES:
            raise ValueError(f"gamma_draw must be one of {GAMMA_DRAW_MODES}; got {gamma_draw!r}")
        self.gamma_draw = gamma_draw
        self.bitpack_output = bitpack_output
        if algo not in ("relay", "plain"):
            raise ValueError(f"Unknown algo '{algo}'. Use 'relay' or 'plain'.")
        self.algo = algo
        perf_norm = (perf or "throughput").strip().lower()
        if perf_norm not in _PERF_ALIASES:
            raise ValueError(f"Unknown perf '{perf_norm}'. Use 'default'/'throughput'.")
        # Both reference modes compute the same per-shot results; one kernel serves both.
        self.perf = _PERF_ALIASES[perf_norm]
        self.alpha_iteration_scaling_factor = float(alpha_iteration_scaling_factor or 1.0)
        scheduling_norm = (scheduling or "flood-fill").strip().lower()
        if scheduling_norm not in ("flood-fill", "layered"):
            raise ValueError(
                f"Unknown scheduling '{scheduling_norm}'. Use 'flood-fill' or 'layered'."
            )
        if scheduling_norm == "layered" and self.perf != "throughput":
            raise ValueError("scheduling='layered' requires perf='throughput'")
        if scheduling_norm == "layered":
            raise NotImplementedError("scheduling='layered' is not implemented on MI355X")
        self.scheduling = scheduling_norm
        self.layered_k_iters = int(layered_k_iters) if layered_k_iters is not None else 0

        self.C, self.V = (int(x) for x in H_csr.shape)
        self._explicit_gammas = None
        if explicit_gammas is not None:
            eg = np.asarray(explicit_gammas, dtype=np.float32)
            if eg.ndim != 2 or eg.shape[1] != self.V:
                raise ValueError(f"explicit_gammas must have shape (K, V={self.V})")
            self._explicit_gammas = torch.from_numpy(eg).to(self.device, dtype=self.dtype)
        self._gamma_by_error = None
        if gamma_dist_interval_by_error is not None:
            gd = np.asarray(gamma_dist_interval_by_error, dtype=np.float32)
            if gd.ndim != 2 or gd.shape[0] != self.V or gd.shape[1] != 2:
                raise ValueError(
                    f"gamma_dist_interval_by_error must have shape (V={self.V}, 2), got "
                    f"{tuple(gd.shape)}"
                )
            if np.any(gd[:, 0] > gd[:, 1]):
                raise ValueError(
                    "gamma_dist_interval_by_error: lower bounds must be <= upper bounds"
                )
            self._gamma_by_error = torch.from_numpy(gd).to(self.device, dtype=self.dtype)

        self.graph = CSRGraph.from_matrix(H_csr)
        self.E = self.graph.E
        log_ratios = compute_log_prior_ratios(error_priors)
        lpr = torch.from_numpy(log_ratios).to(self.device, dtype=self.dtype).unsqueeze(0)
        self.log_prior_ratios_per_gpu = [lpr]
        wj = lpr.squeeze(0)
        self.wj = torch.where(torch.isfinite(wj), wj, torch.zeros_like(wj)).contiguous()
        self.wj_per_gpu = {self.device_ids[0]: self.wj}
        torch.manual_seed(seed)  # the reference's global side effect (decoder.py:774)

        self._block_size = int(block_size) if block_size else 0
        self._placement = -1 if state_placement is None else int(state_placement)
        self._handle = _GraphHandle(self.graph, self.device)
        self._plans: Dict[int, _native.PelPlan] = {}
        if autotune is None:
            autotune = os.environ.get("PELAGOS_AMD_AUTOTUNE", "1") != "0"
        self.autotune = bool(autotune)
        # (B, learned) -> (block, requeue_below, placement)
        self._tuned: Dict[Tuple[int, bool], Tuple[int, int, int]] = {}
        self.tune_runs = 0  # first-sight tunings this decoder timed (cache misses)
        self._workspace: Optional[torch.Tensor] = None
        self._gamma_cache: Dict[Tuple, torch.Tensor] = {}
        self._gamma_perm_cache: Dict[Tuple, Tuple[torch.Tensor, torch.Tensor]] = {}
        n_alpha = max(self.pre_iter, self.set_max_iter, 1)
        self._alpha_cache: Dict[int, torch.Tensor] = {}
        self._alpha_n = n_alpha
        # When True, decode() records per-shot device timestamps in
        # self.shot_clock ([B, 3] int64: start, end in 100 MHz ticks, workgroup).
        self.record_shot_clock = False
        self.shot_clock: Optional[torch.Tensor] = None
        # When a list, decode() appends a (start, end) pair of HIP events
        # recorded on the decode stream around each kernel launch.
        self.kernel_events: Optional[list] = None
        # Re-queue threshold of the launch plan (None: the library's default; 0:
        # every shot runs to completion in one go).
        self.requeue_below: Optional[int] = None
        self.learned_legs = int(learned_legs)
        if self.learned_legs < 0:
            raise ValueError("learned_legs must be >= 0")
        self.learned_last_damp = float(learned_last_damp)
        if not 0.0 <= self.learned_last_damp <= 1.0:
            raise ValueError("learned_last_damp must lie in [0, 1]")
        self.learned_params: Optional[torch.Tensor] = None
        if learned_gamma is not None:
            if self._typed is not None:
                raise NotImplementedError(f"learned_gamma with data_type={self.dtype_str!r}: the network's "
                                          "strengths feed the float32 kernel only")
            prm = getattr(learned_gamma, "params", learned_gamma)
            prm = torch.as_tensor(prm, dtype=torch.float32).reshape(-1)
            if prm.numel() != _native.PEL_MLP_PARAMS:
                raise ValueError(f"learned_gamma must hold {_native.PEL_MLP_PARAMS} parameters")
            self.learned_params = prm.to(self.device).contiguous()
        # When True, decode() with learned strengths keeps the strengths it used
        # in self.learned_gammas ([B, learned_legs, V], caller's variable order).
        self.record_learned_gammas = False
        self.learned_gammas: Optional[torch.Tensor] = None

    # ------------------------------------------------------------------ helpers
    @staticmethod
    def _parse_device_spec(device) -> List[int]:
        """Device list parsing of decoder.py:1403-1452."""
        if isinstance(device, bool):
            raise ValueError(f"Invalid device type: {type(device)}. Expected List[int], int, or str.")
        if isinstance(device, int):
            if device < 1:
                raise ValueError(f"Number of GPUs must be >= 1, got {device}")
            available = torch.cuda.device_count()
            if device > available:
                raise ValueError(f"Requested {device} GPUs but only {available} available")
            return list(range(device))
        if isinstance(device, list):
            if not device:
                raise ValueError("Device list cannot be empty")
            available = torch.cuda.device_count()
            for gpu_id in device:
                if not isinstance(gpu_id, int):
                    raise ValueError(f"Device IDs must be integers, got {type(gpu_id)}")
                if gpu_id < 0 or gpu_id >= available:
                    raise ValueError(f"Device ID {gpu_id} out of range [0, {available - 1}]")
            return device
        if isinstance(device, str):
            if device.strip().lower() in ("auto", "cuda", "rocm"):
                return [0]
            raise ValueError(
                f"Unsupported device string '{device}'. Use List[int], int, or 'auto'/'cuda'."
            )
        raise ValueError(f"Invalid device type: {type(device)}. Expected List[int], int, or str.")

    def gamma_table(self, B: int) -> torch.Tensor:
        """Relay-leg memory strengths: [R, V] (broadcast) or [R, B, V] (per_row).

        Row l is what the reference writes into ``gamma`` at relay leg l
        (decoder.py:3684-3700): explicit row (l+1) % K, else a draw seeded
        ``seed + l`` on the decoding device.
        """
        R, V = self.num_sets, self.V
        per_row = self._explicit_gammas is None and self.gamma_draw == "per_row"
        key = ("per_row", B, R) if per_row else ("broadcast", R)
        cached = self._gamma_cache.get(key)
        if cached is not None:
            return cached
        dev = self.device
        if R == 0:
            table = torch.zeros(1, V, dtype=self.dtype, device=dev)
        elif self._explicit_gammas is not None:
            K = int(self._explicit_gammas.shape[0])
            idx = torch.tensor(explicit_leg_rows(K, R), device=dev)
            table = self._explicit_gammas.index_select(0, idx).contiguous()
        else:
            rows = B if per_row else 1
            draws = []
            for leg in range(R):
                if self._gamma_by_error is not None:
                    g = sample_gamma_per_variable(rows, V, self._gamma_by_error, dev, self.seed + leg)
                else:
                    gmin, gmax = self.gamma_dist_interval
                    g = sample_gamma_uniform(rows, V, gmin, gmax, dev, self.seed + leg)
                draws.append(g.to(self.dtype))
            table = torch.stack(draws, 0).contiguous()
            if not per_row:
                table = table.reshape(R, V)
        if per_row:
            self._gamma_cache = {k: v for k, v in self._gamma_cache.items() if k[0] != "per_row"}
        self._gamma_cache[key] = table
        return table

    def _gamma_by_position(self, table: torch.Tensor) -> torch.Tensor:
        """The [R, V] gamma table in the graph's internal variable order (cached
        per source table; pel_graph_permute_vars)."""
        key = (table.data_ptr(), tuple(table.shape))
        hit = self._gamma_perm_cache.get(key)
        if hit is not None and hit[0] is table:
            return hit[1]
        out = torch.empty_like(table)
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream(self.device).cuda_stream
            _native.check(_native.lib().pel_graph_permute_vars(
                self._handle.h, table.data_ptr(), out.data_ptr(), int(table.shape[0]),
                ctypes.c_void_p(stream)), "pel_graph_permute_vars")
        self._gamma_perm_cache = {key: (table, out)}
        return out

    def _alpha_table(self, n: int) -> torch.Tensor:
        n = max(n, 1)
        t = self._alpha_cache.get(n)
        if t is None:
            a = alpha_schedule(self.normalized_min_sum_alpha, self.alpha_iteration_scaling_factor, n)
            t = torch.from_numpy(a).to(self.device)
            self._alpha_cache[n] = t
        return t

    def _float32_only(self, what: str) -> None:
        if self._typed is not None:
            raise NotImplementedError(f"{what} runs the float32 kernels; data_type={self.dtype_str!r} "
                                      "decodes with decode()")

    def plan(self, B: int, learned: bool = False, legwise: bool = False) -> _native.PelPlan:
        if legwise:
            p = self._plans.get((B, learned, True))
            if p is None:
                p = self._handle.plan(B, self._block_size, self._placement, learned, legwise=True)
                self._plans[(B, learned, True)] = p
            return p
        p = self._plans.get((B, learned))
        if p is None:
            tuned = self._tuned.get((B, learned))
            block = self._block_size or (tuned[0] if tuned else 0)
            place = self._placement if self._placement >= 0 or not tuned else tuned[2]
            p = self._handle.plan(B, block, place, learned)
            if self.requeue_below is None and tuned is not None and (tuned[1] == 0 or p.requeue_below > 0):
                p.requeue_below = int(tuned[1])
            if self.requeue_below is not None:
                if self.requeue_below < 0:
                    raise ValueError("requeue_below must be >= 0")
                if self.requeue_below > 0 and p.requeue_below == 0:
                    raise ValueError("this plan has no workspace for re-queued shots")
                p.requeue_below = int(self.requeue_below)
            self._plans[(B, learned)] = p
        return p

    # ------------------------------------------------------------- autotuning
    def _tune_key(self, B: int, learned: bool) -> str:
        name = torch.cuda.get_device_name(self.device)
        return (f"{self.graph.sparsity_hash}|B={B}|learned={int(learned)}|placement={self._placement}"
                f"|{name}|abi={_native.PEL_ABI_VERSION}")

    def _candidates(self, B: int, learned: bool) -> List[Tuple[int, int, int]]:
        """Launch shapes to time: the block sizes around the library's own choice,
        the re-queue thresholds 0, 16 and 32, and the state placement the library
        picks (the largest that fits in LDS) beside the next smaller one (for
        learned strengths, the fast placements only)."""
        base_plan = self._handle.plan(B, 0, self._placement, learned)
        base = base_plan.block
        blocks = sorted({b for b in (base // 4, base // 2, base) if 64 <= b <= 1024})
        if self._block_size:
            blocks = [self._block_size]
        requeue = [int(self.requeue_below)] if self.requeue_below is not None else [0, 16, 32]
        if self._placement >= 0:
            places = [self._placement]
        else:
            order = [15, 7, 3, 1, 0]
            top = int(base_plan.placement)
            lower = [q for q in order[order.index(top) + 1:] if not learned or (q & 7) == 7]
            places = [top] + lower[:1]
        return [(b, r, q) for q in places for b in blocks for r in requeue]

    def _tune(self, B: int, learned: bool, run) -> None:
        """First sight of (B, learned): the cached choice, or time every
        candidate with ``run()`` (one decode of the caller's syndromes), keep
        the fastest and save it."""
        key = self._tune_key(B, learned)
        self._tuned[(B, learned)] = (self._block_size, 0, self._placement)  # placeholder while timing
        with _TUNE_LOCK:
            table = _load_tune_cache()
        hit = table.get(key)
        if hit is not None:
            self._tuned[(B, learned)] = (int(hit["block"]), int(hit["requeue_below"]), int(hit.get("placement", -1)))
            self._plans.pop((B, learned), None)
            return
        best = None
        for blk, rq, pl in self._candidates(B, learned):
            self._tuned[(B, learned)] = (blk, rq, pl)
            self._plans.pop((B, learned), None)
            try:
                if self.plan(B, learned).requeue_below != rq:
                    continue  # this plan has no workspace for re-queued shots
                run()  # warm-up
                ms = []
                for _ in range(3):
                    ev = []
                    self.kernel_events = ev
                    run()
                    self.kernel_events = None
                    torch.cuda.synchronize(self.device)
                    ms.append(ev[0][0].elapsed_time(ev[0][1]))
            except (ValueError, RuntimeError):
                self.kernel_events = None
                continue  # a shape the graph does not allow
            t = min(ms)
            if best is None or t < best[3]:
                best = (blk, rq, pl, t)
        self.tune_runs += 1
        if best is None:
            del self._tuned[(B, learned)]
            self._plans.pop((B, learned), None)
            return
        self._tuned[(B, learned)] = (best[0], best[1], best[2])
        self._plans.pop((B, learned), None)
        with _TUNE_LOCK:
            table = _load_tune_cache()
            table[key] = {"block": best[0], "requeue_below": best[1], "placement": best[2], "ms": round(best[3], 4)}
            _save_tune_cache(table)

    def _shot_clock_for(self, B: int) -> Optional[torch.Tensor]:
        if not self.record_shot_clock:
            return None
        self.shot_clock = torch.zeros(B, 3, dtype=torch.int64, device=self.device)
        return self.shot_clock

    def _workspace_for(self, nbytes: int) -> torch.Tensor:
        if self._workspace is None or self._workspace.numel() < nbytes:
            self._workspace = torch.empty(int(nbytes), dtype=torch.uint8, device=self.device)
        return self._workspace

    def override_llr_columns(self, cols, value):
        """Write prior LLRs (and selection weights) of ``cols`` directly, bypassing
        the prior clip; returns a restore callable (decoder.py:786-835)."""
        cols = np.asarray(cols, dtype=np.int64)
        lpr = self.log_prior_ratios_per_gpu[0]
        saved_lpr = lpr[0, cols].clone()
        saved_wj = self.wj[cols].clone()
        lpr[0, cols] = torch.as_tensor(value, dtype=lpr.dtype, device=lpr.device)
        self.wj[cols] = torch.as_tensor(value, dtype=self.wj.dtype, device=self.wj.device)

        def _restore():
            lpr[0, cols] = saved_lpr
            self.wj[cols] = saved_wj

        return _restore

    # ------------------------------------------------------------------ decode
    def decode(
        self,
        syndromes: torch.Tensor,
        *,
        init_posterior: Optional[torch.Tensor] = None,
        init_gamma: Optional[torch.Tensor] = None,
        skip_relay_legs: bool = False,
        pre_iter_override: Optional[int] = None,
        return_last_posterior: bool = False,
        prior_llr: Optional[torch.Tensor] = None,
        _records: Optional[dict] = None,
    ) -> Dict[str, torch.Tensor]:
        """Decode ``syndromes [B, C]`` (uint8); returns the reference's result dict.

        Keys: ``errors`` [B,V] uint8 (or [B, ceil(V/32)] int32 when
        bitpack_output), ``weights`` [B] float32 (+inf without a solution),
        ``valid_mask`` [B] bool (found_count >= stop_nconv), ``iterations`` [B]
        int32, ``found_count`` [B] int32, ``unconverged_indices`` [U] int32 and,
        with ``return_last_posterior``, ``last_posterior`` [B,V].
        """
        if syndromes.dim() != 2:
            raise ValueError(f"syndromes must be 2-D [B, C], got shape {tuple(syndromes.shape)}")
        B = int(syndromes.shape[0])
        if syndromes.shape[1] != self.C:
            raise ValueError(f"Expected {self.C} syndrome bits, got {syndromes.shape[1]}")
        V = self.V
        if init_posterior is not None and tuple(init_posterior.shape) != (B, V):
            raise ValueError(f"init_posterior shape {tuple(init_posterior.shape)} != ({B}, {V})")
        if init_gamma is not None and tuple(init_gamma.shape) != (B, V):
            raise ValueError(f"init_gamma shape {tuple(init_gamma.shape)} != ({B}, {V})")
        if prior_llr is not None and tuple(prior_llr.shape) != (B, V):
            raise ValueError(f"prior_llr shape {tuple(prior_llr.shape)} != ({B}, {V})")
        legwise = self.fast_max_unconverged_pct is not None and not skip_relay_legs and _records is None
        dev = self.device
        pre_iter = self.pre_iter if pre_iter_override is None else int(pre_iter_override)
        if pre_iter < 0:
            raise ValueError("pre_iter_override must be >= 0")
        with torch.cuda.device(dev):
            syn = syndromes.to(device=dev, dtype=torch.uint8).contiguous()
            if prior_llr is not None:
                prior = prior_llr.to(device=dev, dtype=self.dtype).contiguous()
                wj = torch.where(torch.isfinite(prior), prior, torch.zeros_like(prior)).contiguous()
            else:
                prior = self.log_prior_ratios_per_gpu[0].reshape(V)
                wj = self.wj
            ig = None
            if init_gamma is not None:
                ig = init_gamma.to(device=dev, dtype=self.dtype).contiguous()
            ip = None
            if init_posterior is not None:
                ip = init_posterior.to(device=dev, dtype=self.dtype).contiguous()
            gtab = self.gamma_table(B)
            alpha = self._alpha_table(max(pre_iter, self.set_max_iter))
            errors = torch.empty(B, V, dtype=torch.uint8, device=dev)
            weights = torch.empty(B, dtype=self.dtype, device=dev)
            iters = torch.empty(B, dtype=torch.int32, device=dev)
            found = torch.empty(B, dtype=torch.int32, device=dev)
            last_post = torch.empty(B, V, dtype=self.dtype, device=dev) if return_last_posterior else None
            learn = _records
            if learn is None and self.learned_params is not None and self.learned_legs > 0:
                learn = dict(mlp=self.learned_params, legs=self.learned_legs)
                if self.record_learned_gammas:
                    self.learned_gammas = torch.zeros(B, self.learned_legs, V, dtype=self.dtype, device=dev)
                    learn["gamma_out"] = self.learned_gammas
            if B > 0 and self._typed is not None:
                if _records is not None:
                    raise NotImplementedError("traces of the learned network are recorded in float32 only")
                typed = (syn, prior, wj, prior_llr is not None, gtab, ig, ip, errors, weights, iters, found,
                         last_post, pre_iter, skip_relay_legs, B)
                if legwise:
                    self._each_leg(B, found, lambda k: self._launch_typed(*typed, leg_step=k))
                else:
                    self._launch_typed(*typed)
            elif B > 0 and legwise:
                self._decode_legwise(syn, prior, wj, prior_llr is not None, gtab, ig, ip, errors, weights,
                                     iters, found, last_post, pre_iter, B, learn)
            elif B > 0:
                args = (syn, prior, wj, prior_llr is not None, gtab, ig, ip, errors, weights,
                        iters, found, last_post, pre_iter, skip_relay_legs, B, learn)
                lkey = (B, learn is not None)
                if self.autotune and lkey not in self._tuned and self.kernel_events is None \
                        and not self.record_shot_clock:
                    self._tune(B, learn is not None, lambda: self._launch(*args))
                self._launch(*args)
            return self._assemble(errors, weights, iters, found, last_post)

    # ------------------------------------------------------------ serving path
    def prepare(self, B: int) -> int:
        """Build everything a decode of B shots needs (launch plan, gamma rows in
        position order, alpha table) so that decode_async makes no allocation
        and no host synchronisation.  Returns the workspace bytes per launch."""
        self._float32_only("prepare")
        learned = self.learned_params is not None and self.learned_legs > 0
        plan = self.plan(B, learned)
        gtab = self.gamma_table(B)
        if gtab.dim() == 2:
            self._gamma_by_position(gtab)
        self._alpha_table(max(self.pre_iter, self.set_max_iter))
        torch.cuda.synchronize(self.device)
        return int(plan.workspace_bytes)

    def decode_async(self, syndromes: torch.Tensor, errors: torch.Tensor, weights: torch.Tensor,
                     iterations: torch.Tensor, found_count: torch.Tensor, workspace: Optional[torch.Tensor],
                     stream: torch.cuda.Stream) -> None:
        """Decode device-resident ``syndromes [B, C]`` into caller-owned buffers
        (errors [B, V] uint8, weights [B] float32, iterations / found_count [B]
        int32) with a caller-owned workspace, ordered on ``stream``, without
        allocating or waiting: the launch a serving loop or a captured graph
        makes.  Call prepare(B) once first (broadcast gamma draws only).  With
        ``workspace=None`` the launch uses the library's own per-GPU workspace
        (``pel_decode_noalloc``, at most PEL_NOALLOC_MAX_BYTES; such calls
        must not overlap on one GPU)."""
        self._float32_only("decode_async")
        B = int(syndromes.shape[0])
        # the buffers are handed to the kernel as raw pointers: check what it assumes
        for name, t, shape, dtype in (("syndromes", syndromes, (B, self.C), torch.uint8),
                                      ("errors", errors, (B, self.V), torch.uint8),
                                      ("weights", weights, (B,), self.dtype),
                                      ("iterations", iterations, (B,), torch.int32),
                                      ("found_count", found_count, (B,), torch.int32)):
            if tuple(t.shape) != shape or t.dtype != dtype or not t.is_contiguous() or t.device != self.device:
                raise ValueError(f"decode_async: {name} must be a contiguous {dtype} tensor of shape {shape} on "
                                 f"{self.device}, got {t.dtype} {tuple(t.shape)} on {t.device}"
                                 f"{'' if t.is_contiguous() else ' (not contiguous)'}")
        if workspace is not None and (workspace.dtype != torch.uint8 or not workspace.is_contiguous()
                                      or workspace.device != self.device):
            raise ValueError("decode_async: workspace must be a contiguous uint8 tensor on the decoder's device")
        if B == 0:
            return
        if self.gamma_draw == "per_row" and self._explicit_gammas is None:
            raise ValueError("decode_async needs broadcast gamma rows (gamma_draw='broadcast')")
        learn = None
        if self.learned_params is not None and self.learned_legs > 0:
            learn = dict(mlp=self.learned_params, legs=self.learned_legs)
        self._launch(syndromes, self.log_prior_ratios_per_gpu[0].reshape(self.V), self.wj, False,
                     self.gamma_table(B), None, None, errors, weights, iterations, found_count, None,
                     self.pre_iter, False, B, learn, workspace=workspace, stream=stream,
                     noalloc=workspace is None)

    def collect_traces(self, syndromes: torch.Tensor, legs: int, with_network: bool = False) -> Dict[str, torch.Tensor]:
        """Decode as ``decode`` does and record, at the start of relay legs
        0..legs-1, the learned network's features of every variable
        (``features`` [B, legs, V, 8]) and the hard decision (``hard``
        [B, legs, V]; 255 where the shot stopped earlier).  The legs run on
        their uniform draws unless ``with_network`` (then this decoder's
        network shifts those of them that start from a valid solution)."""
        B = int(syndromes.shape[0])
        feats = torch.zeros(B, legs, self.V, 8, dtype=torch.float32, device=self.device)
        hard = torch.full((B, legs, self.V), 255, dtype=torch.uint8, device=self.device)
        mlp = None
        if with_network:
            if self.learned_params is None:
                raise ValueError("collect_traces(with_network=True) needs learned_gamma")
            mlp = self.learned_params
        out = self.decode(syndromes, _records=dict(mlp=mlp, legs=int(legs), feat_out=feats, hard_out=hard))
        out["features"] = feats
        out["hard"] = hard
        return out

    def _launch(self, syn, prior, wj, prior_batched, gtab, ig, ip, errors, weights, iters, found,
                last_post, pre_iter, skip_relay_legs, B, learn=None, workspace=None, stream=None,
                noalloc=False, leg_step=0) -> None:
        plan = self.plan(B, learn is not None, legwise=leg_step > 0)
        if noalloc:
            ws = None
        elif workspace is not None:
            if workspace.numel() < plan.workspace_bytes:
                raise ValueError(f"workspace holds {workspace.numel()} bytes, the plan needs {plan.workspace_bytes}")
            ws = workspace
        else:
            ws = self._workspace_for(plan.workspace_bytes)
        params, batch = self._args(syn, prior, wj, prior_batched, gtab, ig, ip, errors, weights, iters, found,
                                   last_post, pre_iter, skip_relay_legs, B, learn, ws)
        params.leg_step = int(leg_step)
        cur = stream if stream is not None else torch.cuda.current_stream(self.device)
        ev = self.kernel_events
        if ev is not None:
            # The library records the pair right around the decode kernel on this
            # stream; recording them once here creates the underlying HIP events.
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(cur)
            e1.record(cur)
            batch.ev_begin = e0.cuda_event
            batch.ev_end = e1.cuda_event
        fn = _native.lib().pel_decode_noalloc if noalloc else _native.lib().pel_decode
        _native.check(fn(self._handle.h, ctypes.byref(params), ctypes.byref(batch), ctypes.byref(plan),
                         ctypes.c_void_p(cur.cuda_stream)), "pel_decode_noalloc" if noalloc else "pel_decode")
        if ev is not None:
            ev.append((e0, e1))

    def _launch_typed(self, syn, prior, wj, prior_batched, gtab, ig, ip, errors, weights, iters, found,
                      last_post, pre_iter, skip_relay_legs, B, leg_step: int = 0) -> None:
        """A float16 / float64 decode (pel_decode_typed): buffers of this decoder's dtype, gamma rows in
        the caller's variable order; leg_step k > 0 runs leg k - 1 only (as _launch)."""
        lib = _native.lib()
        grid = ctypes.c_int32()
        nbytes = ctypes.c_int64()
        flags = _native.PEL_PLAN_LEGWISE if leg_step > 0 else 0
        _native.check(lib.pel_plan_typed(self._handle.h, B, self._typed, flags, ctypes.byref(grid),
                                         ctypes.byref(nbytes)), "pel_plan_typed")
        ws = self._workspace_for(int(nbytes.value))
        params, batch = self._args(syn, prior, wj, prior_batched, gtab, ig, ip, errors, weights, iters, found,
                                   last_post, pre_iter, skip_relay_legs, B, None, ws, by_position=False)
        params.leg_step = int(leg_step)
        cur = torch.cuda.current_stream(self.device)
        ev = self.kernel_events
        if ev is not None:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(cur)
            e1.record(cur)
            batch.ev_begin = e0.cuda_event
            batch.ev_end = e1.cuda_event
        _native.check(lib.pel_decode_typed(self._handle.h, ctypes.byref(params), ctypes.byref(batch), self._typed,
                                           grid.value, ctypes.c_void_p(cur.cuda_stream)), "pel_decode_typed")
        if ev is not None:
            ev.append((e0, e1))

    def _args(self, syn, prior, wj, prior_batched, gtab, ig, ip, errors, weights, iters, found, last_post,
              pre_iter, skip_relay_legs, B, learn, ws, by_position: bool = True):
        """The (pel_params, pel_batch) pair of one decode."""
        per_row = gtab.dim() == 3
        # the kernel reads rows 0 .. num_sets-1 (and, per row, B shots) of the table
        if (self.num_sets > 0 and gtab.shape[0] < self.num_sets) or (per_row and gtab.shape[1] != B) \
                or gtab.shape[-1] != self.V:
            raise ValueError(f"gamma table of shape {tuple(gtab.shape)} does not cover num_sets={self.num_sets}, "
                             f"B={B}, V={self.V}")
        if not per_row and by_position:
            gtab = self._gamma_by_position(gtab)
        params = _native.PelParams(
            pre_iter=pre_iter,
            num_sets=self.num_sets,
            set_max_iter=self.set_max_iter,
            stop_nconv=self.stop_nconv,
            plain=int(self.algo == "plain"),
            gamma0=float(self.gamma0),
            use_alpha=int(self.normalized_min_sum_alpha is not None),
            use_beta=int(self.offset_min_sum_beta is not None),
            beta=float(self.offset_min_sum_beta or 0.0),
            alpha_t=self._alpha_table(max(pre_iter, self.set_max_iter)).data_ptr(),
            last_leg=int(self.unconverged_correction == "last_leg"),
            skip_relay_legs=int(bool(skip_relay_legs)),
        )
        batch = _native.PelBatch(
            B=B,
            syndromes=syn.data_ptr(),
            prior_llr=prior.data_ptr(),
            wj=wj.data_ptr(),
            prior_batched=int(prior_batched),
            gamma_legs=gtab.data_ptr(),
            gamma_per_row=int(per_row),
            gamma_by_position=int(not per_row and by_position),
            init_gamma=_ptr(ig),
            init_posterior=_ptr(ip),
            errors=errors.data_ptr(),
            weights=weights.data_ptr(),
            iterations=iters.data_ptr(),
            found_count=found.data_ptr(),
            last_posterior=_ptr(last_post),
            workspace=None if ws is None else ws.data_ptr(),
            shot_clock=_ptr(self._shot_clock_for(B)) if B > 0 else None,
        )
        if learn is not None:
            batch.mlp = _ptr(learn.get("mlp"))
            batch.learned_legs = int(learn["legs"])
            batch.feat_out = _ptr(learn.get("feat_out"))
            batch.hard_out = _ptr(learn.get("hard_out"))
            batch.gamma_out = _ptr(learn.get("gamma_out"))
            batch.learned_lo = float(self.gamma_dist_interval[0])
            batch.learned_hi = float(self.gamma_dist_interval[1])
            batch.learned_last_damp = float(learn.get("last_damp", self.learned_last_damp))
        return params, batch

    def job_args(self, syndromes: torch.Tensor, errors: torch.Tensor, weights: torch.Tensor,
                 iterations: torch.Tensor, found_count: torch.Tensor, workspace: torch.Tensor):
        """(pel_params, pel_batch) of one job of a multi-job launch (``pel_decode_jobs``):
        a decode of device-resident ``syndromes`` into caller-owned buffers, as
        ``decode_async`` makes it.  Call prepare(B) first."""
        self._float32_only("job_args")
        B = int(syndromes.shape[0])
        learn = None
        if self.learned_params is not None and self.learned_legs > 0:
            learn = dict(mlp=self.learned_params, legs=self.learned_legs)
        return self._args(syndromes, self.log_prior_ratios_per_gpu[0].reshape(self.V), self.wj, False,
                          self.gamma_table(B), None, None, errors, weights, iterations, found_count, None,
                          self.pre_iter, False, B, learn, workspace)

    def _assemble(self, errors, weights, iters, found, last_post) -> Dict[str, torch.Tensor]:
        unconverged = torch.nonzero(found < self.stop_nconv, as_tuple=False).flatten().to(torch.int32)
        out = {
            "errors": bitpack_errors(errors) if self.bitpack_output else errors,
            "weights": weights,
            "valid_mask": found >= self.stop_nconv,
            "iterations": iters,
            "found_count": found,
            "unconverged_indices": unconverged,
        }
        if last_post is not None:
            out["last_posterior"] = last_post
        return out

    def _decode_legwise(self, syn, prior, wj, prior_batched, gtab, ig, ip, errors, weights, iters, found,
                        last_post, pre_iter, B, learn) -> None:
        """The reference's leg-synchronous schedule with its fast_max_unconverged_pct
        exit (decoder.py:2890-2937, 3655-3835): one launch per leg (the pre-iter
        leg, then relay legs 0, 1, ...), each running that leg for the shots still
        short of stop_nconv solutions from the state the previous launch parked and
        leaving every shot's outputs as if the decode ended there.  After each leg
        the found counts are read (one host synchronisation per leg, as the
        reference's): the decode stops when every shot has stop_nconv solutions,
        or when the fraction without them is below fast_max_unconverged_pct (and
        below 1)."""
        self._each_leg(B, found, lambda k: self._launch(syn, prior, wj, prior_batched, gtab, ig, ip, errors,
                                                        weights, iters, found, last_post, pre_iter, False, B,
                                                        learn, leg_step=k))

    def _each_leg(self, B, found, launch) -> None:
        """Drive leg-stepped launches launch(1), launch(2), ... (the pre-iter leg, then relay legs
        0, 1, ...) until the fast_max_unconverged_pct exit on the found counts (decoder.py:2890-2937)
        or the last leg."""
        pct = float(self.fast_max_unconverged_pct)
        for k in range(self.num_sets + 1):
            launch(k + 1)
            if k == self.num_sets:
                break
            unconverged = int((found < self.stop_nconv).sum()) / B
            if unconverged == 0.0 or (unconverged < pct and unconverged < 1.0):
                break

    def get_stats(self) -> Dict:
        return {
            "graph": self.graph.get_stats(),
            "parameters": {
                "pre_iter": self.pre_iter,
                "num_sets": self.num_sets,
                "set_max_iter": self.set_max_iter,
                "stop_nconv": self.stop_nconv,
                "gamma0": self.gamma0,
                "gamma_dist_interval": self.gamma_dist_interval,
                "algo": self.algo,
            },
        }
