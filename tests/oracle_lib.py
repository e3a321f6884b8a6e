"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: it is the checker the HIP path is compared
against, never a fallback for it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
LIB_PATH = ROOT / "oracle" / "liboracle.so"

u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


class ScaleInfo(C.Structure):
    _fields_ = [("orig_w", C.c_int), ("orig_h", C.c_int), ("work_w", C.c_int), ("work_h", C.c_int),
                ("sx", C.c_double), ("sy", C.c_double), ("D", C.c_double * 9)]


class Top2(C.Structure):
    _fields_ = [("first_distance", C.c_double), ("second_distance", C.c_double),
                ("first_index", C.c_int), ("second_index", C.c_int)]


class Point(C.Structure):
    _fields_ = [("ax", C.c_double), ("ay", C.c_double), ("bx", C.c_double), ("by", C.c_double)]


class HomographyResult(C.Structure):
    _fields_ = [("H", C.c_double * 9), ("H_s", C.c_double * 9), ("num_matches", C.c_int), ("num_inliers", C.c_int),
                ("inlier_ratio", C.c_double), ("inlier_coverage", C.c_double), ("reproj_residual", C.c_double),
                ("status", C.c_int)]


class Model(C.Structure):
    _fields_ = [("h", C.c_double * 9), ("loss", C.c_double), ("hypothesis", C.c_int), ("valid", C.c_int)]


class S5State(C.Structure):
    _fields_ = [("best", Model), ("evaluated", C.c_int), ("status", C.c_int)]


class Geometry(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("origin_x", C.c_int), ("origin_y", C.c_int),
                ("width_a", C.c_int), ("height_a", C.c_int), ("width_b", C.c_int), ("height_b", C.c_int),
                ("denominator_tolerance", C.c_double), ("snap_tolerance", C.c_double),
                ("inverse", C.c_double * 9)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            subprocess.run(["make", "-s", "-C", str(LIB_PATH.parent)], check=True)
        L = C.CDLL(str(LIB_PATH))
        L.orc_make_scale_info.argtypes = [C.c_int, C.c_int, C.c_double, C.POINTER(ScaleInfo)]
        L.orc_s1_preprocess.argtypes = [u8p, C.c_int, C.c_int, C.c_int, C.c_double, u8p, C.c_int,
                                        C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_s1_gradient.argtypes = [u8p, C.c_int, C.c_int, C.c_int]
        L.orc_s1_gradient.restype = C.c_double
        L.orc_fast_score.argtypes = [u8p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_harris_response.argtypes = [u8p, C.c_int, C.c_int, C.c_int, C.c_float]
        L.orc_harris_response.restype = C.c_float
        L.orc_s2_detect.argtypes = [u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                                    f32p, f32p, f32p, C.c_void_p, C.c_void_p]
        L.orc_s3_blur.argtypes = [u8p, C.c_int, C.c_int, C.c_int, u8p]
        L.orc_fast_atan2.argtypes = [C.c_float, C.c_float]
        L.orc_fast_atan2.restype = C.c_float
        L.orc_s3_describe.argtypes = [u8p, C.c_int, C.c_int, C.c_int, f32p, f32p, C.c_int, C.c_void_p, u8p]
        L.orc_s4_top2.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                  C.POINTER(Top2)]
        L.orc_s4_match.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                   C.c_double, C.c_int, i32p, i32p, f32p, C.c_void_p]
        L.orc_s5_random_word.argtypes = [C.c_uint64]
        L.orc_s5_random_word.restype = C.c_uint64
        L.orc_s5_sample_four.argtypes = [C.c_int, C.c_uint64, C.c_int, i32p]
        L.orc_s5_nondegenerate.argtypes = [C.POINTER(Point)]
        L.orc_s5_fit.argtypes = [C.c_int, C.POINTER(Point), C.c_void_p, C.POINTER(C.c_double)]
        L.orc_s5_normalize.argtypes = [C.POINTER(C.c_double)]
        L.orc_s5_residual.argtypes = [C.POINTER(C.c_double), Point]
        L.orc_s5_residual.restype = C.c_double
        L.orc_s5_valid_domain.argtypes = [C.POINTER(C.c_double), C.c_int, C.c_int]
        L.orc_s5_tree_loss.argtypes = [C.c_int, C.POINTER(Point), C.POINTER(C.c_double), C.c_double]
        L.orc_s5_tree_loss.restype = C.c_double
        L.orc_s5_serial_loss.argtypes = [C.c_int, C.POINTER(Point), C.POINTER(C.c_double), C.c_double]
        L.orc_s5_serial_loss.restype = C.c_double
        L.orc_s5_estimate.argtypes = [C.c_int, C.POINTER(Point), C.POINTER(ScaleInfo), C.POINTER(ScaleInfo),
                                      C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64, C.POINTER(S5State),
                                      u8p, C.POINTER(HomographyResult)]
        L.orc_s5_estimate_kp.argtypes = [f32p, f32p, C.c_int, C.c_int, f32p, f32p, C.c_int, C.c_int, i32p, i32p,
                                         C.c_int, C.c_int, C.POINTER(ScaleInfo), C.POINTER(ScaleInfo), C.c_int,
                                         C.c_int, C.c_int, C.c_double, C.c_uint64, C.POINTER(S5State), u8p,
                                         C.POINTER(HomographyResult)]
        L.orc_s6_prepare_geometry.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_double,
                                              C.c_double, C.c_double, C.c_ulonglong, C.POINTER(Geometry)]
        L.orc_s6_effective_reductions.argtypes = [C.c_int, C.c_int, C.c_int]
        L.orc_s6_render.argtypes = [u8p, C.c_int, u8p, C.c_int, C.POINTER(Geometry), C.c_int, u8p, u8p,
                                    C.POINTER(C.c_ulonglong)]
        L.orc_s6_prepare_chain.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double),
                                           C.c_double, C.c_double, C.c_double, C.c_ulonglong, C.POINTER(Geometry)]
        L.orc_s6_render_chain.argtypes = [C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_int), C.POINTER(Geometry),
                                          C.c_int, u8p, u8p, C.POINTER(C.c_ulonglong)]
        vp = C.c_void_p
        L.orc_scene_features.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, C.c_int, C.c_int, C.c_int,
                                         C.c_int, C.c_int, vp]
        L.orc_gate_features.argtypes = ([vp, C.c_int, C.c_int, C.c_int, C.c_int] * 2 +
                                        [vp, vp, C.c_int, vp, vp, C.c_int, vp, vp, vp, C.c_int, vp, C.c_int,
                                         C.c_int, C.c_double, C.c_double, C.c_double, vp, C.POINTER(ScaleInfo),
                                         C.POINTER(ScaleInfo), vp, C.c_uint, vp])
        _lib = L
    return _lib


# ---------------------------------------------------------------- wrappers

def scale_info(w, h, s):
    si = ScaleInfo()
    st = lib().orc_make_scale_info(w, h, s, C.byref(si))
    return st, si


def s1_preprocess(rgb: np.ndarray, scale: float) -> np.ndarray:
    h, w, _ = rgb.shape
    st, si = scale_info(w, h, scale)
    assert st == 0
    out = np.zeros((si.work_h, si.work_w), np.uint8)
    ww, wh = C.c_int(), C.c_int()
    assert lib().orc_s1_preprocess(np.ascontiguousarray(rgb), w, h, 3 * w, scale, out, si.work_w,
                                   C.byref(ww), C.byref(wh)) == 0
    return out


def s1_gradient(gray: np.ndarray) -> float:
    h, w = gray.shape
    return lib().orc_s1_gradient(np.ascontiguousarray(gray), w, h, w)


def s2_detect(gray: np.ndarray, fast_threshold: int, k: int, harris_k: float = 0.04, maps=False):
    h, w = gray.shape
    xs = np.zeros(max(k, 1), np.float32)
    ys = np.zeros_like(xs)
    rs = np.zeros_like(xs)
    scores = np.zeros((h, w), np.uint8) if maps else None
    maxima = np.zeros((h, w), np.uint8) if maps else None
    n = lib().orc_s2_detect(np.ascontiguousarray(gray), w, h, w, fast_threshold, k, harris_k, xs, ys, rs,
                            scores.ctypes.data if maps else None, maxima.ctypes.data if maps else None)
    assert n >= 0
    kp = np.stack([xs[:n], ys[:n], rs[:n]], axis=1)
    return (kp, scores, maxima) if maps else kp


def s3_blur(gray: np.ndarray) -> np.ndarray:
    h, w = gray.shape
    out = np.zeros((h, w), np.uint8)
    lib().orc_s3_blur(np.ascontiguousarray(gray), w, h, w, out)
    return out


def s3_describe(gray: np.ndarray, kp: np.ndarray) -> np.ndarray:
    h, w = gray.shape
    n = kp.shape[0]
    out = np.zeros((max(n, 1), 32), np.uint8)
    xs = np.ascontiguousarray(kp[:, 0], np.float32) if n else np.zeros(1, np.float32)
    ys = np.ascontiguousarray(kp[:, 1], np.float32) if n else np.zeros(1, np.float32)
    lib().orc_s3_describe(np.ascontiguousarray(gray), w, h, w, xs, ys, n, None, out)
    return out[:n]


def s4_top2(da: np.ndarray, db: np.ndarray):
    binary = da.dtype == np.uint8
    na, dim = da.shape
    nb = db.shape[0]
    esz = 1 if binary else 4
    out = (Top2 * max(na, 1))()
    da = np.ascontiguousarray(da)
    db = np.ascontiguousarray(db)
    lib().orc_s4_top2(int(binary), da.ctypes.data, na, dim * esz, db.ctypes.data, nb, dim * esz, dim, out)
    return [(out[i].first_distance, out[i].second_distance, out[i].first_index, out[i].second_index)
            for i in range(na)]


def s4_match(da: np.ndarray, db: np.ndarray, ratio: float, cross_check: bool):
    binary = da.dtype == np.uint8
    na, dim = da.shape
    nb = db.shape[0]
    esz = 1 if binary else 4
    ia = np.zeros(max(na, 1), np.int32)
    ib = np.zeros_like(ia)
    dist = np.zeros(max(na, 1), np.float32)
    da = np.ascontiguousarray(da)
    db = np.ascontiguousarray(db)
    n = lib().orc_s4_match(int(binary), da.ctypes.data, na, dim * esz, db.ctypes.data, nb, dim * esz, dim,
                           ratio, int(cross_check), ia, ib, dist, None)
    return ia[:n], ib[:n], dist[:n]


def points(arr: np.ndarray):
    """arr (n, 4) float64 [ax, ay, bx, by] -> ctypes Point array."""
    n = arr.shape[0]
    pts = (Point * max(n, 1))()
    for i in range(n):
        pts[i] = Point(*map(float, arr[i]))
    return pts


def s5_estimate_kp(ax, ay, bx, by, idx_a, idx_b, sa, sb, hypotheses, refinement_steps, threshold, seed,
                   previous_hypotheses=0, state=None, counts=None, caps=None):
    ax, ay, bx, by = (np.ascontiguousarray(v, np.float32) for v in (ax, ay, bx, by))
    idx_a = np.ascontiguousarray(idx_a, np.int32)
    idx_b = np.ascontiguousarray(idx_b, np.int32)
    na, nb, nm = counts if counts else (len(ax), len(bx), len(idx_a))
    ca, cb, cm = caps if caps else (len(ax), len(bx), len(idx_a))
    state = state if state is not None else S5State()
    mask = np.zeros(max(cm, 1), np.uint8)
    res = HomographyResult()
    lib().orc_s5_estimate_kp(ax if len(ax) else np.zeros(1, np.float32), ay if len(ay) else np.zeros(1, np.float32),
                             na, ca, bx if len(bx) else np.zeros(1, np.float32),
                             by if len(by) else np.zeros(1, np.float32), nb, cb,
                             idx_a if len(idx_a) else np.zeros(1, np.int32),
                             idx_b if len(idx_b) else np.zeros(1, np.int32), nm, cm, C.byref(sa), C.byref(sb),
                             hypotheses, previous_hypotheses, refinement_steps, threshold, seed, C.byref(state),
                             mask, C.byref(res))
    return res, mask[:cm], state


def prepare_geometry(wa, ha, wb, hb, h, den_tol=1e-12, inv_tol=1e-12, snap_tol=1e-10, max_pixels=16777216):
    g = Geometry()
    harr = (C.c_double * 9)(*[float(v) for v in np.asarray(h, np.float64).ravel()])
    st = lib().orc_s6_prepare_geometry(wa, ha, wb, hb, harr, den_tol, inv_tol, snap_tol, max_pixels, C.byref(g))
    return st, g


MODES = {"feather": 0, "multiband_3": 1, "multiband_5": 2, "multiband_6": 3, "multiband_7": 4}


def prepare_chain(sizes, hs, max_pixels=16777216, den_tol=1e-12, inv_tol=1e-12, snap_tol=1e-10):
    """Canvas of a chain: sizes [(w, h)] of n frames, hs the n - 1 adjacent
    homographies (hs[k]: frame k + 1 -> frame k). (status, [Geometry] * n)."""
    n = len(sizes)
    w = (C.c_int * n)(*[s[0] for s in sizes])
    h = (C.c_int * n)(*[s[1] for s in sizes])
    flat = [float(v) for m in hs for v in np.asarray(m, np.float64).ravel()]
    harr = (C.c_double * max(len(flat), 1))(*flat)
    out = (Geometry * n)()
    st = lib().orc_s6_prepare_chain(n, w, h, harr, den_tol, inv_tol, snap_tol, max_pixels, out)
    return st, list(out)


def s6_render_chain(frames, geoms, mode):
    """frames: n (h, w, 3) uint8 arrays; geoms from prepare_chain."""
    n = len(frames)
    frames = [np.ascontiguousarray(f) for f in frames]
    g0 = geoms[0]
    out = np.zeros((g0.height, g0.width, 3), np.uint8)
    mask = np.zeros((g0.height, g0.width), np.uint8)
    stats = (C.c_ulonglong * 2)()
    ptrs = (C.c_void_p * n)(*[f.ctypes.data for f in frames])
    pitches = (C.c_int * n)(*[f.shape[1] * 3 for f in frames])
    garr = (Geometry * n)(*[Geometry.from_buffer_copy(bytes(g)) for g in geoms])  # any same-layout struct
    st = lib().orc_s6_render_chain(n, ptrs, pitches, garr, MODES[mode] if isinstance(mode, str) else mode, out, mask,
                                   stats)
    assert st == 0
    return out, mask, (stats[0], stats[1])


def s6_render(a: np.ndarray, b: np.ndarray, g: Geometry, mode: int):
    out = np.zeros((g.height, g.width, 3), np.uint8)
    mask = np.zeros((g.height, g.width), np.uint8)
    stats = (C.c_ulonglong * 2)()
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    st = lib().orc_s6_render(a, a.shape[1] * 3, b, b.shape[1] * 3, C.byref(g), mode, out, mask, stats)
    assert st == 0
    return out, mask, (stats[0], stats[1])


def _image(im: np.ndarray):
    """(h, w) GRAY8 or (h, w, 3) RGB8 -> (array, w, h, pitch, format)."""
    im = np.ascontiguousarray(im, np.uint8)
    return im, im.shape[1], im.shape[0], im.strides[0], 1 if im.ndim == 2 else 0


def scene_features(gray_a: np.ndarray, gray_b: np.ndarray, relations: bool = True):
    """(status, 32 doubles) of features.cu:125-135."""
    a, wa, ha, pa, fa = _image(gray_a)
    b, wb, hb, pb, fb = _image(gray_b)
    out = np.zeros(32)
    st = lib().orc_scene_features(a.ctypes.data, wa, ha, pa, fa, b.ctypes.data, wb, hb, pb, fb, int(relations),
                                  out.ctypes.data)
    return st, out


def gate_features(img_a, img_b, kp_a, kp_b, idx_a, idx_b, distance, mask, num_matches, num_inliers,
                  inlier_ratio, inlier_coverage, reproj_residual, h, sa, sb, forward=None, relation_mask=7,
                  capacity=None):
    """(status, 38 doubles) of features.cu:136-170. kp_a / kp_b: (capacity, 2)
    float32 keypoint x, y arrays (rows up to the keypoint capacity); forward:
    a list of (first_distance, second_distance, first_index, second_index)
    per row of kp_a, or None."""
    a, wa, ha, pa, fa = _image(img_a)
    b, wb, hb, pb, fb = _image(img_b)
    kax = np.ascontiguousarray(kp_a[:, 0], np.float32)
    kay = np.ascontiguousarray(kp_a[:, 1], np.float32)
    kbx = np.ascontiguousarray(kp_b[:, 0], np.float32)
    kby = np.ascontiguousarray(kp_b[:, 1], np.float32)
    ia = np.ascontiguousarray(idx_a, np.int32)
    ib = np.ascontiguousarray(idx_b, np.int32)
    dist = np.ascontiguousarray(distance, np.float32)
    m = np.ascontiguousarray(mask, np.uint8)
    hm = np.ascontiguousarray(h, np.float64).reshape(9)
    fw = None
    if forward is not None:
        fw = (Top2 * max(len(forward), 1))()
        for i, t in enumerate(forward):
            fw[i] = Top2(*t)
    out = np.zeros(38)
    st = lib().orc_gate_features(a.ctypes.data, wa, ha, pa, fa, b.ctypes.data, wb, hb, pb, fb,
                                 kax.ctypes.data, kay.ctypes.data, len(kax), kbx.ctypes.data, kby.ctypes.data,
                                 len(kbx), ia.ctypes.data, ib.ctypes.data, dist.ctypes.data,
                                 len(ia) if capacity is None else capacity, m.ctypes.data, num_matches,
                                 num_inliers, inlier_ratio, inlier_coverage, reproj_residual, hm.ctypes.data,
                                 C.byref(sa), C.byref(sb), C.cast(fw, C.c_void_p) if fw is not None else None,
                                 relation_mask, out.ctypes.data)
    return st, out
