"""Python mirror of the harness API: PipelineRunner (harness/pipeline_runner.hpp:32-81),
HarnessConfig / RunRecord (harness/harness_types.hpp:14-106), the legacy-v1
controller (harness/controller_policy.hpp:208-226) and the seeded synthetic
pair generator used by the benchmarks."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi as F
from .stages import StageError, _check, image

TEXTURED, LOW_TEXTURE, REPETITIVE = 0, 1, 2


def synth_pair(seed: int, width: int, height: int, overlap=(0.3, 0.5), kind: int = TEXTURED):
    """(A, B, H_ba): uint8 (H, W, 3) images and the true homography B -> A."""
    a = np.empty((height, width, 3), np.uint8)
    b = np.empty_like(a)
    h = (C.c_double * 9)()
    _check(F.load().cs_synth_pair(seed, width, height, overlap[0], overlap[1], kind, a.ctypes.data, b.ctypes.data, h),
           "synth_pair")
    return a, b, np.array(h[:]).reshape(3, 3)


def synth_strip(seed: int, n: int, width: int, height: int, overlap: float = 0.4, kind: int = TEXTURED):
    """(frames, h_adj): n uint8 (H, W, 3) frames and the n - 1 true homographies
    frame k + 1 -> frame k."""
    frames = np.empty((n, height, width, 3), np.uint8)
    h = (C.c_double * max(1, 9 * (n - 1)))()
    _check(F.load().cs_synth_strip(seed, n, width, height, overlap, kind, frames.ctypes.data, h), "synth_strip")
    return frames, [np.array(h[9 * k: 9 * k + 9]).reshape(3, 3) for k in range(n - 1)]


def harness_config(S=1.0, K=1000, R=256, blend="feather", fast_threshold=20, harris_k=0.04, ratio=0.8,
                   cross_check=False, batch=64, refinements=1, threshold=3.0, seed=0, pyramid_levels=1,
                   pyramid_scale_factor=1.2) -> F.HarnessConfig:
    mode = F.BLEND_NAMES[blend] if isinstance(blend, str) else int(blend)
    return F.HarnessConfig(S, K, R, mode, fast_threshold, harris_k, ratio, int(cross_check), batch, refinements,
                           threshold, seed, pyramid_levels, pyramid_scale_factor)


def record_dict(r: F.RunRecord) -> dict:
    d = {name: getattr(r, name) for name, _ in F.RunRecord._fields_}
    d["H"] = list(r.H)
    d["status"] = r.status.decode()
    d["fail_stage"] = r.fail_stage.decode()
    return d


class Controller:
    """The built-in legacy-v1 ladder (blind: legacy-v1-blind-s05), or a
    controller-policy-v2/v3/v4 document given as policy_json."""

    def __init__(self, blind: bool = False, policy_json: str | None = None):
        if policy_json is None:
            self.handle = F.load().cs_controller_create_legacy(int(blind))
            if not self.handle:
                raise RuntimeError("controller policy rejected")
        else:
            err = C.create_string_buffer(512)
            self.handle = F.load().cs_controller_create_policy(policy_json.encode(), err, len(err))
            if not self.handle:
                raise ValueError(f"controller policy rejected: {err.value.decode()}")

    def observations(self):
        """(uses_learned_features, uses_pair_relations, relation_mask, needs_gate_features)."""
        a, b, d = C.c_int(), C.c_int(), C.c_int()
        m = C.c_uint()
        F.load().cs_controller_observations(self.handle, C.byref(a), C.byref(b), C.byref(m), C.byref(d))
        return bool(a.value), bool(b.value), m.value, bool(d.value)

    def walk(self, grad_a, grad_b, wa, ha, wb, hb, feedback, exec_status, scene=None, gates=None):
        """begin_pair then observe over the feedbacks; scene: 32 doubles or
        None; gates: per feedback 38 doubles or None. Rows are (action,
        reason, S, K, R), the first the initial knobs."""
        n = len(feedback)
        fb = (F.HomographyResult * max(n, 1))(*feedback)
        st = (C.c_int * max(n, 1))(*exec_status)
        trace = (C.c_double * (5 * (n + 2)))()
        sc = np.ascontiguousarray(scene, np.float64) if scene is not None else None
        g = np.zeros((max(n, 1), F.GATE_FEATURES))
        present = (C.c_int * max(n, 1))()
        for i, row in enumerate(gates or []):
            if row is not None:
                g[i] = row
                present[i] = 1
        rows = F.load().cs_controller_walk_features(self.handle, grad_a, grad_b, wa, ha, wb, hb,
                                                    sc.ctypes.data if sc is not None else None, fb, st,
                                                    g.ctypes.data, present, n, trace)
        return [tuple(trace[5 * i:5 * i + 5]) for i in range(rows)]

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                F.load().cs_controller_destroy(self.handle)
                self.handle = None
        except (AttributeError, TypeError):  # interpreter teardown: the library module is already gone
            pass


class PipelineRunner:
    def __init__(self, max_width, max_height, max_k=4000, max_r=2048, canvas_width=4096, canvas_height=4096,
                 workspace_mode=F.MULTIBAND5, device_reserve=256 << 20):
        L = F.load()
        self.handle = L.cs_runner_create()
        if not self.handle:
            raise RuntimeError("cs_runner_create failed")
        _check(L.cs_runner_initialize(self.handle, max_width, max_height, max_k, max_r, canvas_width, canvas_height,
                                      workspace_mode, device_reserve), "runner.initialize")

    def upload_pair(self, a: np.ndarray, b: np.ndarray):
        a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
        _check(F.load().cs_runner_upload_pair(self.handle, a.ctypes.data, a.shape[1], a.shape[0], b.ctypes.data,
                                              b.shape[1], b.shape[0]), "runner.upload_pair")

    def use_device_pair(self, a, b):
        ia, ib = image(a), image(b)
        _check(F.load().cs_runner_use_device_pair(self.handle, C.byref(ia), C.byref(ib)), "runner.use_device_pair")

    def execute_run(self, cfg: F.HarnessConfig) -> F.RunRecord:
        rec = F.RunRecord()
        _check(F.load().cs_runner_execute_run(self.handle, C.byref(cfg), C.byref(rec)), "runner.execute_run")
        return rec

    def execute_adaptive_run(self, controller: Controller, seed: int = 0) -> F.RunRecord:
        rec = F.RunRecord()
        _check(F.load().cs_runner_execute_adaptive_run(self.handle, controller.handle, seed, C.byref(rec)),
               "runner.execute_adaptive_run")
        return rec

    def attempt_trace(self) -> list:
        """Per-attempt records of the last adaptive run (knobs, decision, S5 result, stage times)."""
        n = C.c_int()
        out = (F.AttemptRecord * 16)()
        _check(F.load().cs_runner_attempt_trace(self.handle, out, 16, C.byref(n)), "runner.attempt_trace")
        return list(out[: n.value])

    def set_flags(self, reuse=True, probe=True):
        F.load().cs_runner_set_flags(self.handle, int(reuse), int(probe))

    def set_collect_features(self, enabled=True):
        """Scene features once each pair is on the device, gate features after
        each successful S5 of execute_run (PipelineRunner::set_collect_features)."""
        _check(F.load().cs_runner_set_collect_features(self.handle, int(enabled)), "runner.set_collect_features")

    def features(self):
        """(scene[32], gate[38], features_present, t_features) of the last run."""
        scene = np.zeros(F.SCENE_FEATURES)
        gate = np.zeros(F.GATE_FEATURES)
        present, t = C.c_int(), C.c_double()
        _check(F.load().cs_runner_features(self.handle, scene.ctypes.data, gate.ctypes.data, C.byref(present),
                                           C.byref(t)), "runner.features")
        return scene, gate, bool(present.value), t.value

    def outputs(self):
        """Device views of the last run's outputs (valid until the next run)."""
        canvas, mask = F.Image(), F.Image()
        ka, kb = F.Keypoints(), F.Keypoints()
        da, db = F.Descriptors(), F.Descriptors()
        m = F.Matches()
        inl = C.c_void_p()
        h = F.HomographyResult()
        F.load().cs_runner_outputs(self.handle, C.byref(canvas), C.byref(mask), C.byref(ka), C.byref(kb),
                                   C.byref(da), C.byref(db), C.byref(m), C.byref(inl), C.byref(h))
        return dict(canvas=canvas, mask=mask, ka=ka, kb=kb, da=da, db=db, matches=m, inlier_mask=inl.value, h=h)

    def close(self):
        if getattr(self, "handle", None):
            F.load().cs_runner_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except (AttributeError, TypeError):  # interpreter teardown: the library module is already gone
            pass


class StripStitcher:
    """n frames in a row onto one canvas (cs_strip_*): S1-S5 of every adjacent
    pair in one batched phase, one chain render. Frames are device RGB8 tensors."""

    def __init__(self, max_frames, max_width, max_height, max_k=16000, max_r=8192):
        h = C.c_void_p()
        _check(F.load().cs_strip_create(max_frames, max_width, max_height, max_k, max_r, C.byref(h)), "strip.create")
        self.handle = h.value

    def run(self, frames, cfg: F.HarnessConfig):
        """(status, record, [HomographyResult] * (n - 1), keypoint counts)."""
        n = len(frames)
        views = (F.Image * n)(*[image(f) for f in frames])
        rec = F.StripRecord()
        pairs = (F.HomographyResult * (n - 1))()
        counts = (C.c_int * n)()
        st = F.load().cs_strip_run(self.handle, n, views, C.byref(cfg), C.byref(rec), pairs, counts)
        return st, rec, list(pairs), list(counts)

    def outputs(self, n):
        """Device views of the last canvas and mask, and the n frame geometries."""
        canvas, mask = F.Image(), F.Image()
        g = (F.Geometry * n)()
        _check(F.load().cs_strip_outputs(self.handle, C.byref(canvas), C.byref(mask), g, n), "strip.outputs")
        return canvas, mask, list(g)

    def close(self):
        if getattr(self, "handle", None):
            F.load().cs_strip_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except (AttributeError, TypeError):
            pass


_hip = None


def _hip_memcpy_d2h(dst: np.ndarray, ptr: int, nbytes: int) -> None:
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemcpy.restype = C.c_int
    if nbytes:
        rc = _hip.hipMemcpy(dst.ctypes.data, C.c_void_p(ptr), nbytes, 2)  # hipMemcpyDeviceToHost
        if rc != 0:
            raise StageError("hipMemcpy", F.DEVICE_ERROR)


def download(ptr: int, nbytes: int) -> np.ndarray:
    """Synchronous copy of nbytes of device memory to a host array."""
    out = np.empty(max(nbytes, 1), np.uint8)
    _hip_memcpy_d2h(out, ptr, nbytes)
    return out[:nbytes]


def download_image(img: F.Image, channels: int) -> np.ndarray:
    """Synchronous copy of a pitched device image to a host (H, W[, 3]) array."""
    raw = download(img.data, img.pitch_bytes * (img.height - 1) + img.width * channels)
    raw = np.concatenate([raw, np.zeros(img.pitch_bytes * img.height - raw.size, np.uint8)])
    arr = raw.reshape(img.height, img.pitch_bytes)[:, : img.width * channels]
    return arr.reshape(img.height, img.width, channels).copy() if channels == 3 else arr.copy()
