"""Python mirror of the stage APIs (S1..S6) over the C ABI.

Argument meaning, status codes and ownership follow the reference's stage
headers (src/s1/s1.hpp, src/s2_s3/s2.hpp, src/s2_s3/s3.hpp, src/s4/s4.hpp,
src/s5/s5.hpp, src/s6/s6.hpp): the caller owns device buffers, workspaces
are queried then supplied, work is enqueued on the current stream and
counts stay on the device. torch tensors are used only as device memory.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _ffi as F


class StageError(RuntimeError):
    def __init__(self, what: str, status: int):
        super().__init__(f"{what} failed: {F.load().cs_status_name(status).decode()} ({status})")
        self.status = status


def _check(status: int, what: str) -> None:
    if status != F.OK:
        raise StageError(what, status)


def stream_handle(stream: torch.cuda.Stream | None = None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def image(t: torch.Tensor, fmt: int | None = None) -> F.Image:
    """View of a uint8 device tensor (H, W, 3) RGB or (H, W) gray; rows may be pitched."""
    assert t.dtype == torch.uint8 and t.is_cuda
    if t.dim() == 3:
        assert t.shape[2] == 3 and t.stride(2) == 1 and t.stride(1) == 3
        return F.Image(t.data_ptr(), t.shape[1], t.shape[0], t.stride(0), F.RGB8 if fmt is None else fmt)
    assert t.dim() == 2 and t.stride(1) == 1
    return F.Image(t.data_ptr(), t.shape[1], t.shape[0], t.stride(0), F.GRAY8 if fmt is None else fmt)


def pitched(h: int, w: int, channels: int, pad: int = 64, device="cuda") -> torch.Tensor:
    """Device image with row pitch w*channels + pad (the reference harness pads rows by 64 B)."""
    base = torch.zeros((h, w * channels + pad), dtype=torch.uint8, device=device)
    v = base[:, : w * channels]
    return v.view(h, w, channels) if channels == 3 else v


def workspace(nbytes: int, device="cuda") -> torch.Tensor:
    """256-byte aligned device workspace (torch's allocator aligns to at least 512 B)."""
    t = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)
    assert t.data_ptr() % 256 == 0
    return t


def scale_info(w: int, h: int, s: float) -> F.ScaleInfo:
    si = F.ScaleInfo()
    _check(F.load().cs_make_scale_info(w, h, s, C.byref(si)), "make_scale_info")
    return si


# ---------------------------------------------------------------- S1

def s1_preprocess(a: torch.Tensor, b: torch.Tensor, scale: float, pad: int = 64):
    L = F.load()
    sa, sb = scale_info(a.shape[1], a.shape[0], scale), scale_info(b.shape[1], b.shape[0], scale)
    ga, gb = pitched(sa.work_h, sa.work_w, 1, pad), pitched(sb.work_h, sb.work_w, 1, pad)
    ia, ib, oa, ob = image(a), image(b), image(ga), image(gb)
    _check(L.cs_s1_preprocess(C.byref(ia), C.byref(ib), scale, C.byref(oa), C.byref(ob), C.byref(sa), C.byref(sb),
                              stream_handle()), "s1_preprocess")
    return ga, gb, sa, sb


def s1_preprocess_single(a: torch.Tensor, scale: float, pad: int = 64):
    L = F.load()
    si = scale_info(a.shape[1], a.shape[0], scale)
    g = pitched(si.work_h, si.work_w, 1, pad)
    ia, og = image(a), image(g)
    _check(L.cs_s1_preprocess_single(C.byref(ia), scale, C.byref(og), C.byref(si), stream_handle()),
           "s1_preprocess_single")
    return g, si


def s1_compute_gradient(ga: torch.Tensor, gb: torch.Tensor):
    L = F.load()
    ws = workspace(32)
    pinned = torch.zeros(2, dtype=torch.float64, pin_memory=True)
    ia, ib = image(ga), image(gb)
    _check(L.cs_s1_compute_gradient(C.byref(ia), C.byref(ib), ws.data_ptr(), pinned.data_ptr(),
                                    pinned.data_ptr() + 8, stream_handle()), "s1_compute_gradient")
    torch.cuda.current_stream().synchronize()
    return float(pinned[0]), float(pinned[1])


# ---------------------------------------------------------------- S2 / S3

@dataclass
class KeypointSet:
    x: torch.Tensor
    y: torch.Tensor
    response: torch.Tensor | None
    count: torch.Tensor  # device int32[1]

    @property
    def capacity(self) -> int:
        return self.x.numel()

    def c(self, capacity: int | None = None) -> F.Keypoints:
        cap = self.capacity if capacity is None else capacity
        return F.Keypoints(self.x.data_ptr(), self.y.data_ptr(),
                           self.response.data_ptr() if self.response is not None else None, cap,
                           self.count.data_ptr())

    @staticmethod
    def empty(capacity: int, with_response: bool = True) -> "KeypointSet":
        n = max(capacity, 1)
        return KeypointSet(torch.zeros(n, dtype=torch.float32, device="cuda"),
                           torch.zeros(n, dtype=torch.float32, device="cuda"),
                           torch.zeros(n, dtype=torch.float32, device="cuda") if with_response else None,
                           torch.zeros(1, dtype=torch.int32, device="cuda"))

    def host(self):
        n = int(self.count.item())
        return n, self.x[:n].cpu(), self.y[:n].cpu(), (self.response[:n].cpu() if self.response is not None else None)


def s2_detect(gray: torch.Tensor, fast_threshold: int, max_keypoints: int, harris_k: float = 0.04,
              out: KeypointSet | None = None, ws: torch.Tensor | None = None) -> KeypointSet:
    L = F.load()
    cfg = F.S2Config(fast_threshold, max_keypoints, harris_k)
    nbytes = C.c_size_t()
    _check(L.cs_s2_workspace_bytes(gray.shape[1], gray.shape[0], C.byref(cfg), C.byref(nbytes)), "s2_workspace_bytes")
    ws = ws if ws is not None else workspace(nbytes.value)
    out = out if out is not None else KeypointSet.empty(max_keypoints)
    img, kp = image(gray), out.c(max_keypoints)
    _check(L.cs_s2_detect(C.byref(img), C.byref(cfg), C.byref(kp), ws.data_ptr(), ws.numel(), stream_handle()),
           "s2_detect")
    return out


def s3_describe(gray: torch.Tensor, kp: KeypointSet, capacity: int | None = None, stride: int = 32,
                ws: torch.Tensor | None = None) -> torch.Tensor:
    L = F.load()
    cap = kp.capacity if capacity is None else capacity
    nbytes = C.c_size_t()
    _check(L.cs_s3_workspace_bytes(gray.shape[1], gray.shape[0], C.byref(nbytes)), "s3_workspace_bytes")
    ws = ws if ws is not None else workspace(nbytes.value)
    desc = torch.zeros((max(cap, 1), stride), dtype=torch.uint8, device="cuda")
    d = F.Descriptors(F.BINARY_U8, 32, stride, desc.data_ptr())
    img, k = image(gray), kp.c(cap)
    _check(L.cs_s3_describe(C.byref(img), C.byref(k), C.byref(d), ws.data_ptr(), ws.numel(), stream_handle()),
           "s3_describe")
    return desc


# ---------------------------------------------------------------- S4

def descriptors(t: torch.Tensor) -> F.Descriptors:
    """(rows, dim) uint8 (binary) or float32 device tensor with unit column stride."""
    assert t.is_cuda and t.dim() == 2 and t.stride(1) == 1
    if t.dtype == torch.uint8:
        return F.Descriptors(F.BINARY_U8, t.shape[1], t.stride(0), t.data_ptr())
    assert t.dtype == torch.float32
    return F.Descriptors(F.FLOAT32, t.shape[1], t.stride(0) * 4, t.data_ptr())


@dataclass
class MatchSet:
    idx_a: torch.Tensor
    idx_b: torch.Tensor
    distance: torch.Tensor
    count: torch.Tensor

    def c(self) -> F.Matches:
        return F.Matches(self.idx_a.data_ptr(), self.idx_b.data_ptr(), self.distance.data_ptr(), self.idx_a.numel(),
                         self.count.data_ptr())

    @staticmethod
    def empty(capacity: int) -> "MatchSet":
        n = max(capacity, 1)
        return MatchSet(torch.full((n,), -7, dtype=torch.int32, device="cuda"),
                        torch.full((n,), -7, dtype=torch.int32, device="cuda"),
                        torch.zeros(n, dtype=torch.float32, device="cuda"),
                        torch.zeros(1, dtype=torch.int32, device="cuda"))

    def host(self):
        n = int(self.count.item())
        return self.idx_a[:n].cpu(), self.idx_b[:n].cpu(), self.distance[:n].cpu()


def _kp_view(count: torch.Tensor, capacity: int) -> F.Keypoints:
    return F.Keypoints(None, None, None, capacity, count.data_ptr())


def s4_match(da: torch.Tensor, count_a: torch.Tensor, db: torch.Tensor, count_b: torch.Tensor, ratio: float,
             cross_check: bool, capacity_a: int | None = None, capacity_b: int | None = None,
             out: MatchSet | None = None, ws: torch.Tensor | None = None):
    L = F.load()
    ca = da.shape[0] if capacity_a is None else capacity_a
    cb = db.shape[0] if capacity_b is None else capacity_b
    cfg = F.S4Config(ratio, int(cross_check))
    nbytes = C.c_size_t()
    _check(L.cs_s4_workspace_bytes(ca, cb, C.byref(cfg), C.byref(nbytes)), "s4_workspace_bytes")
    ws = ws if ws is not None else workspace(nbytes.value)
    out = out if out is not None else MatchSet.empty(ca)
    ka, kb = _kp_view(count_a, ca), _kp_view(count_b, cb)
    ea, eb, m = descriptors(da), descriptors(db), out.c()
    _check(L.cs_s4_match(C.byref(ka), C.byref(ea), C.byref(kb), C.byref(eb), C.byref(cfg), C.byref(m),
                         ws.data_ptr(), ws.numel(), stream_handle()), "s4_match")
    return out, ws


def s4_knn_top2(da: torch.Tensor, count_a: torch.Tensor, db: torch.Tensor, count_b: torch.Tensor):
    L = F.load()
    ca, cb = da.shape[0], db.shape[0]
    out = torch.zeros((max(ca, 1), 3), dtype=torch.float64, device="cuda")  # 24-byte Top2 records
    ka, kb = _kp_view(count_a, ca), _kp_view(count_b, cb)
    ea, eb = descriptors(da), descriptors(db)
    _check(L.cs_s4_knn_top2(C.byref(ka), C.byref(ea), C.byref(kb), C.byref(eb), out.data_ptr(), stream_handle()),
           "s4_knn_top2")
    return out


def top2_records(t: torch.Tensor, n: int):
    """Decodes n 24-byte Top2 records from a device tensor to Python tuples."""
    raw = t.reshape(-1)[: 3 * n].cpu().numpy().tobytes()
    arr = (F.Top2 * n).from_buffer_copy(raw)
    return [(r.first_distance, r.second_distance, r.first_index, r.second_index) for r in arr]


# ---------------------------------------------------------------- S5

def s5_workspace_bytes(hypotheses: int, previous: int = 0, batch_size: int = 64, refinement_steps: int = 1,
                       threshold: float = 3.0, seed: int = 0, generation: int = 1) -> int:
    cfg = F.S5Config(hypotheses, previous, batch_size, refinement_steps, threshold, seed, generation)
    n = C.c_size_t()
    _check(F.load().cs_s5_workspace_bytes(C.byref(cfg), C.byref(n)), "s5_workspace_bytes")
    return n.value


class S5Runner:
    """Holds one S5 workspace, mask and pinned result so R extensions can reuse state."""

    def __init__(self, max_hypotheses: int, match_capacity: int):
        self.ws = workspace(s5_workspace_bytes(max_hypotheses))
        self.mask = torch.zeros(max(match_capacity, 1), dtype=torch.uint8, device="cuda")
        self.pinned = torch.zeros(C.sizeof(F.HomographyResult), dtype=torch.uint8, pin_memory=True)

    def estimate(self, kp_a: KeypointSet, kp_b: KeypointSet, matches: MatchSet, sa: F.ScaleInfo, sb: F.ScaleInfo,
                 hypotheses: int, previous: int = 0, batch_size: int = 64, refinement_steps: int = 1,
                 threshold: float = 3.0, seed: int = 0, generation: int = 1, capacity_a=None, capacity_b=None):
        L = F.load()
        cfg = F.S5Config(hypotheses, previous, batch_size, refinement_steps, threshold, seed, generation)
        a, b, m = kp_a.c(capacity_a), kp_b.c(capacity_b), matches.c()
        _check(L.cs_s5_estimate(C.byref(a), C.byref(b), C.byref(m), C.byref(sa), C.byref(sb), C.byref(cfg),
                                self.mask.data_ptr(), self.pinned.data_ptr(), self.ws.data_ptr(), self.ws.numel(),
                                stream_handle()), "s5_estimate")
        torch.cuda.current_stream().synchronize()
        return F.HomographyResult.from_buffer_copy(self.pinned.numpy().tobytes())


# ---------------------------------------------------------------- S6

def prepare_geometry(wa: int, ha: int, wb: int, hb: int, h, policy: F.GeometryPolicy | None = None):
    g = F.Geometry()
    harr = (C.c_double * 9)(*[float(v) for v in (h.ravel() if hasattr(h, "ravel") else h)])
    st = F.load().cs_s6_prepare_geometry(wa, ha, wb, hb, harr, C.byref(policy) if policy else None, C.byref(g))
    return st, g


def s6_render(a: torch.Tensor, b: torch.Tensor, g: F.Geometry, mode: int, stats: bool = True, pad: int = 64):
    L = F.load()
    n = C.c_size_t()
    _check(L.cs_s6_workspace_bytes(C.byref(g), mode, C.byref(n)), "s6_workspace_bytes")
    ws = workspace(n.value) if n.value else None
    out = pitched(g.height, g.width, 3, pad)
    mask = pitched(g.height, g.width, 1, pad)
    counters = torch.zeros(2, dtype=torch.int64, device="cuda") if stats else None
    ia, ib, io, im = image(a), image(b), image(out), image(mask)
    _check(L.cs_s6_render(C.byref(ia), C.byref(ib), C.byref(g), mode, C.byref(io), C.byref(im),
                          ws.data_ptr() if ws is not None else None, n.value, stream_handle(),
                          C.cast(counters.data_ptr(), C.POINTER(F.RenderStatistics)) if stats else None),
           "s6_render")
    return out, mask, counters


def prepare_chain(sizes, h_adj, policy: F.GeometryPolicy | None = None):
    """Canvas of a chain of frames: sizes [(w, h)] per frame, h_adj the n - 1
    homographies frame k + 1 -> frame k. Returns (status, [Geometry] * n)."""
    n = len(sizes)
    g = (F.Geometry * n)()
    ws = (C.c_int * n)(*[int(w) for w, _ in sizes])
    hs = (C.c_int * n)(*[int(h) for _, h in sizes])
    flat = [float(v) for h in h_adj for v in (h.ravel() if hasattr(h, "ravel") else h)]
    harr = (C.c_double * max(1, len(flat)))(*flat)
    st = F.load().cs_s6_prepare_chain(n, ws, hs, harr, C.byref(policy) if policy else None, g)
    return st, list(g)


def s6_render_chain(frames, geoms, mode: int, stats: bool = True, pad: int = 64):
    """N-frame render of device RGB8 frames onto the chain canvas: (out, mask, counters)."""
    L = F.load()
    n = len(frames)
    g = (F.Geometry * n)(*geoms)
    nb = C.c_size_t()
    _check(L.cs_s6_chain_workspace_bytes(n, g, mode, C.byref(nb)), "s6_chain_workspace_bytes")
    ws = workspace(nb.value)
    out = pitched(g[0].height, g[0].width, 3, pad)
    mask = pitched(g[0].height, g[0].width, 1, pad)
    counters = torch.zeros(2, dtype=torch.int64, device="cuda") if stats else None
    views = (F.Image * n)(*[image(f) for f in frames])
    io, im = image(out), image(mask)
    _check(L.cs_s6_render_chain(n, views, g, mode, C.byref(io), C.byref(im), ws.data_ptr(), nb.value,
                                stream_handle(),
                                C.cast(counters.data_ptr(), C.POINTER(F.RenderStatistics)) if stats else None),
           "s6_render_chain")
    return out, mask, counters


# ---------------------------------------------------------------- controller observations

def extract_scene(ga: torch.Tensor, gb: torch.Tensor, relations: bool = True):
    """Scene features of two GRAY8 device images (controller::extract_scene,
    src/controller/features.cu:125-135): (status, 32 doubles or None)."""
    L = F.load()
    dev = torch.zeros(F.FEATURE_DEVICE_DOUBLES, dtype=torch.float64, device="cuda")
    host = torch.zeros(F.GATE_FEATURES, dtype=torch.float64, pin_memory=True)
    ia, ib = image(ga), image(gb)
    st = L.cs_extract_scene(C.byref(ia), C.byref(ib), dev.data_ptr(), host.data_ptr(), stream_handle(),
                            int(relations))
    torch.cuda.current_stream().synchronize()
    return st, (host[: F.SCENE_FEATURES].numpy().copy() if st == F.OK else None)


def extract_gate(a: torch.Tensor, b: torch.Tensor, kp_a: F.Keypoints, kp_b: F.Keypoints, matches: F.Matches,
                 mask_ptr: int, h: F.HomographyResult, sa: F.ScaleInfo, sb: F.ScaleInfo, forward_ptr: int | None,
                 relation_mask: int = 7):
    """Gate features of a finished S5 (controller::extract_gate,
    src/controller/features.cu:136-170): (status, 38 doubles or None).
    Keypoints, matches, the inlier mask and the forward top-2 view are
    device views (C structs / pointers)."""
    L = F.load()
    dev = torch.zeros(F.FEATURE_DEVICE_DOUBLES, dtype=torch.float64, device="cuda")
    host = torch.zeros(F.GATE_FEATURES, dtype=torch.float64, pin_memory=True)
    ia, ib = image(a), image(b)
    st = L.cs_extract_gate(C.byref(ia), C.byref(ib), C.byref(kp_a), C.byref(kp_b), C.byref(matches), mask_ptr,
                           C.byref(h), C.byref(sa), C.byref(sb), dev.data_ptr(), host.data_ptr(), stream_handle(),
                           forward_ptr, relation_mask)
    torch.cuda.current_stream().synchronize()
    return st, (host.numpy().copy() if st == F.OK else None)


# ---------------------------------------------------------------- multi-level detection

def pyramid_level_counts(levels: int, scale_factor: float, max_keypoints: int) -> list[int]:
    cfg = F.PyramidConfig(levels, scale_factor, 20, max_keypoints, 0.04)
    out = (C.c_int * max(levels, 1))()
    _check(F.load().cs_pyramid_level_counts(C.byref(cfg), out), "pyramid_level_counts")
    return list(out[:levels])


def pyramid_detect_describe(rgb: torch.Tensor, scale: float, levels: int, scale_factor: float, fast_threshold: int,
                            max_keypoints: int, harris_k: float = 0.04):
    """Keypoints (level-0 work coordinates), 32-byte descriptors and the level
    of each keypoint, over an S1 pyramid of the device RGB image."""
    L = F.load()
    cfg = F.PyramidConfig(levels, scale_factor, fast_threshold, max_keypoints, harris_k)
    n = C.c_size_t()
    _check(L.cs_pyramid_workspace_bytes(rgb.shape[1], rgb.shape[0], scale, C.byref(cfg), C.byref(n)),
           "pyramid_workspace_bytes")
    ws = workspace(n.value)
    kp = KeypointSet.empty(max_keypoints)
    desc = torch.zeros((max(max_keypoints, 1), 32), dtype=torch.uint8, device="cuda")
    lv = torch.zeros(max(max_keypoints, 1), dtype=torch.uint8, device="cuda")
    si = F.ScaleInfo()
    img, k, d = image(rgb), kp.c(max_keypoints), F.Descriptors(F.BINARY_U8, 32, 32, desc.data_ptr())
    _check(L.cs_pyramid_detect_describe(C.byref(img), scale, C.byref(cfg), C.byref(k), C.byref(d), lv.data_ptr(),
                                        C.byref(si), ws.data_ptr(), n.value, stream_handle()), "pyramid_detect_describe")
    return kp, desc, lv, si
