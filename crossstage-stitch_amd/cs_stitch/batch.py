"""Batched and served forms over the C ABI (cs_batch_*, cs_server_*).

BatchRunner runs a ragged batch of pairs (mixed resolutions, keypoint caps,
hypothesis counts and blend modes) with one persistent, graph-replayed launch
per stage kernel; Server replays a request arrival trace with continuous
batching, HBM-budget admission and the adaptive controller per request.
Images are device RGB8 tensors (see stages.image / stages.pitched).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _ffi as F
from .stages import _check, image


def pair_spec(a, b, S=1.0, K=1000, R=256, blend="multiband_5", fast_threshold=20, harris_k=0.04, ratio=0.8,
              cross_check=False, batch_size=64, refinements=1, threshold=3.0, seed=0) -> F.PairSpec:
    mode = F.BLEND_NAMES[blend] if isinstance(blend, str) else int(blend)
    ia = a if isinstance(a, F.Image) else image(a)
    ib = b if isinstance(b, F.Image) else image(b)
    return F.PairSpec(ia, ib, S, K, R, mode, fast_threshold, harris_k, ratio, int(cross_check), batch_size,
                      refinements, threshold, seed)


class BatchRunner:
    def __init__(self, max_pairs: int = 64, pool_bytes: int = 8 << 30, graphs: bool = True):
        self.handle = C.c_void_p()
        _check(F.load().cs_batch_create(max_pairs, pool_bytes, int(graphs), C.byref(self.handle)), "batch_create")
        self.max_pairs = max_pairs
        self._pending = None  # size of the submitted, not yet completed batch

    def run(self, specs):
        """Runs one batch; returns (outcomes, timings). Outcome device views stay valid until the next run."""
        n = len(specs)
        arr = (F.PairSpec * max(n, 1))(*specs)
        out = (F.PairOutcome * max(n, 1))()
        t = F.BatchTimings()
        _check(F.load().cs_batch_run(self.handle, arr, n, out, C.byref(t)), "batch_run")
        self._pending = None
        return list(out[:n]), t

    def submit(self, specs):
        """Stream form: enqueues the batch, whose S1-S5 overlap the previous batch's S6.

        Returns (outcomes of the previous batch, or None when none was pending; timings).
        A batch's device views stay valid until the second submit after its own.
        """
        n = len(specs)
        arr = (F.PairSpec * max(n, 1))(*specs)
        t = F.BatchTimings()
        had = self._pending is not None
        prev_n = self._pending or 0
        prev = (F.PairOutcome * max(prev_n, 1))()
        _check(F.load().cs_batch_submit(self.handle, arr, n, prev, prev_n, C.byref(t)), "batch_submit")
        self._pending = n
        return (list(prev[:prev_n]) if had else None), t

    def finish(self):
        """Completes the last submitted batch; returns its outcomes ([] when none is pending)."""
        n = self._pending or 0
        out = (F.PairOutcome * max(n, 1))()
        _check(F.load().cs_batch_finish(self.handle, out, n), "batch_finish")
        self._pending = None
        return list(out[:n])

    def autotune(self, specs, cache_path: str | None = None) -> F.Tuning:
        arr = (F.PairSpec * len(specs))(*specs)
        t = F.Tuning()
        _check(F.load().cs_batch_autotune(self.handle, arr, len(specs),
                                          cache_path.encode() if cache_path else None, C.byref(t)), "batch_autotune")
        return t

    def apply(self, s4_segment=128, workgroups_per_cu=8, graphs=True, stage_workgroups=(0, 0, 0, 0, 0, 0),
              s1_tile_rows=64):
        """Launch shapes; stage_workgroups overrides workgroups_per_cu per stage S1..S6 where non-zero."""
        t = F.Tuning(s4_segment, workgroups_per_cu, int(graphs), (C.c_int * 6)(*stage_workgroups), s1_tile_rows)
        _check(F.load().cs_batch_apply(self.handle, C.byref(t)), "batch_apply")

    def set_probe(self, on: bool):
        """Per-kernel S6 timing: direct launches with HIP events around each kernel."""
        _check(F.load().cs_batch_set_probe(self.handle, int(on)), "batch_set_probe")

    def probe_read(self):
        """[{kernel, level, launches, total_ms, algorithmic_bytes}] summed over the probed renders."""
        n = C.c_int()
        _check(F.load().cs_batch_probe_read(self.handle, None, 0, C.byref(n)), "batch_probe_read")
        out = (F.KernelProbe * max(n.value, 1))()
        _check(F.load().cs_batch_probe_read(self.handle, out, n.value, C.byref(n)), "batch_probe_read")
        return [dict(kernel=e.kernel.decode(), level=e.level, launches=e.launches, total_ms=e.total_ms,
                     algorithmic_bytes=e.algorithmic_bytes) for e in out[:n.value]]

    def pool(self):
        cap, use, peak = C.c_uint64(), C.c_uint64(), C.c_uint64()
        F.load().cs_batch_pool(self.handle, C.byref(cap), C.byref(use), C.byref(peak))
        return cap.value, use.value, peak.value

    def close(self):
        if getattr(self, "handle", None) and self.handle.value:
            F.load().cs_batch_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except (AttributeError, TypeError):  # interpreter teardown: the library module is already gone
            pass


def pair_bytes(width, height, K, R) -> int:
    """Pool bytes one pair takes for S1-S5 (width/height: max over A and B)."""
    v = C.c_uint64()
    _check(F.load().cs_batch_pair_bytes(width, height, K, R, C.byref(v)), "batch_pair_bytes")
    return v.value


def render_bytes(canvas_width, canvas_height, blend) -> int:
    """Pool bytes one canvas with its S6 workspace takes."""
    mode = F.BLEND_NAMES[blend] if isinstance(blend, str) else int(blend)
    v = C.c_uint64()
    _check(F.load().cs_batch_render_bytes(canvas_width, canvas_height, mode, C.byref(v)), "batch_render_bytes")
    return v.value


def request(id, a, b, arrival_ms, seed=0, fixed: F.PairSpec | None = None, keep_output=False) -> F.Request:
    ia = a if isinstance(a, F.Image) else image(a)
    ib = b if isinstance(b, F.Image) else image(b)
    return F.Request(id, ia, ib, arrival_ms, seed, fixed if fixed is not None else F.PairSpec(), int(keep_output))


def poisson_arrivals(rate_per_s: float, n: int, seed: int) -> list[float]:
    """Arrival times (ms) of a Poisson process: seeded exponential gaps."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1000.0 / rate_per_s, size=n)
    return np.cumsum(gaps).tolist()


class Server:
    def __init__(self, hbm_budget: int = 64 << 30, max_batch: int = 64, max_live: int = 256, adaptive: bool = True,
                 graphs: bool = True, canvas_factor: float = 4.0, policy_json: str | None = None):
        """policy_json: a controller-policy-v2/v3/v4 document for the adaptive
        controller (None: the built-in legacy-v1 ladder)."""
        self.handle = C.c_void_p()
        cfg = F.ServerConfig(hbm_budget, max_batch, max_live, int(adaptive), int(graphs), canvas_factor)
        _check(F.load().cs_server_create_policy(C.byref(cfg), policy_json.encode() if policy_json else None,
                                                C.byref(self.handle)), "server_create")

    def run_trace(self, requests):
        """Replays the trace (sorted by arrival); returns (completions in completion order, stats)."""
        n = len(requests)
        arr = (F.Request * max(n, 1))(*requests)
        out = (F.Completion * max(n, 1))()
        st = F.ServeStats()
        _check(F.load().cs_server_run_trace(self.handle, arr, n, out, C.byref(st)), "server_run_trace")
        return list(out[:n]), st

    def release(self, id):
        F.load().cs_server_release(self.handle, id)

    def close(self):
        if getattr(self, "handle", None) and self.handle.value:
            F.load().cs_server_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except (AttributeError, TypeError):  # interpreter teardown: the library module is already gone
            pass


def percentile(values, q):
    v = sorted(values)
    if not v:
        return 0.0
    pos = q * (len(v) - 1)
    lo = math.floor(pos)
    hi = min(lo + 1, len(v) - 1)
    return v[lo] + (v[hi] - v[lo]) * (pos - lo)
