"""ctypes declarations of include/cs_stitch.h and the loader of libcsstitch.so.

The library is the product: if it is missing, loading raises. There is no
fallback implementation anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libcsstitch.so"

OK, INVALID_ARGUMENT, INSUFFICIENT_WORKSPACE, DEVICE_ERROR, NOT_ENOUGH_POINTS, DEGENERATE = range(6)
RGB8, GRAY8 = 0, 1
# Fixed device array behind cs_s4_match_noalloc (CS_S4_NOALLOC_SCRATCH_BYTES).
S4_NOALLOC_SCRATCH_BYTES = 36 << 20
BINARY_U8, FLOAT32 = 0, 1
FEATHER, MULTIBAND3, MULTIBAND5, MULTIBAND6, MULTIBAND7 = range(5)
BLEND_NAMES = {"feather": FEATHER, "multiband_3": MULTIBAND3, "multiband_5": MULTIBAND5,
               "multiband_6": MULTIBAND6, "multiband_7": MULTIBAND7}


class Image(C.Structure):
    _fields_ = [("data", C.c_void_p), ("width", C.c_int), ("height", C.c_int), ("pitch_bytes", C.c_int),
                ("format", C.c_int)]


class ScaleInfo(C.Structure):
    _fields_ = [("orig_w", C.c_int), ("orig_h", C.c_int), ("work_w", C.c_int), ("work_h", C.c_int),
                ("sx", C.c_double), ("sy", C.c_double), ("D", C.c_double * 9)]


class Keypoints(C.Structure):
    _fields_ = [("x", C.c_void_p), ("y", C.c_void_p), ("response", C.c_void_p), ("capacity", C.c_int),
                ("d_count", C.c_void_p)]


class Descriptors(C.Structure):
    _fields_ = [("type", C.c_int), ("dim", C.c_int), ("row_stride_bytes", C.c_int), ("data", C.c_void_p)]


class Matches(C.Structure):
    _fields_ = [("idx_a", C.c_void_p), ("idx_b", C.c_void_p), ("distance", C.c_void_p), ("capacity", C.c_int),
                ("d_count", C.c_void_p)]


class HomographyResult(C.Structure):
    _fields_ = [("H", C.c_double * 9), ("H_s", C.c_double * 9), ("num_matches", C.c_int), ("num_inliers", C.c_int),
                ("inlier_ratio", C.c_double), ("inlier_coverage", C.c_double), ("reproj_residual", C.c_double),
                ("status", C.c_int)]


class Top2(C.Structure):
    _fields_ = [("first_distance", C.c_double), ("second_distance", C.c_double), ("first_index", C.c_int),
                ("second_index", C.c_int)]


class S2Config(C.Structure):
    _fields_ = [("fast_threshold", C.c_int), ("max_keypoints", C.c_int), ("harris_k", C.c_float)]


class S4Config(C.Structure):
    _fields_ = [("ratio_threshold", C.c_double), ("cross_check", C.c_int)]


class S5Config(C.Structure):
    _fields_ = [("hypotheses", C.c_int), ("previous_hypotheses", C.c_int), ("batch_size", C.c_int),
                ("refinement_steps", C.c_int), ("threshold", C.c_double), ("seed", C.c_uint64),
                ("generation", C.c_uint64)]


class GeometryPolicy(C.Structure):
    _fields_ = [("denominator_tolerance", C.c_double), ("inverse_tolerance", C.c_double),
                ("snap_tolerance", C.c_double), ("max_canvas_pixels", C.c_uint64)]


class Geometry(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("origin_x", C.c_int), ("origin_y", C.c_int),
                ("width_a", C.c_int), ("height_a", C.c_int), ("width_b", C.c_int), ("height_b", C.c_int),
                ("denominator_tolerance", C.c_double), ("snap_tolerance", C.c_double), ("inverse", C.c_double * 9)]


class RenderStatistics(C.Structure):
    _fields_ = [("union_pixels", C.c_uint64), ("overlap_pixels", C.c_uint64)]


class PyramidConfig(C.Structure):
    _fields_ = [("levels", C.c_int), ("scale_factor", C.c_double), ("fast_threshold", C.c_int),
                ("max_keypoints", C.c_int), ("harris_k", C.c_float)]


class HarnessConfig(C.Structure):
    _fields_ = [("S", C.c_double), ("K", C.c_int), ("R", C.c_int), ("blend_mode", C.c_int),
                ("fast_threshold", C.c_int), ("harris_k", C.c_float), ("ratio_threshold", C.c_double),
                ("cross_check", C.c_int), ("msac_batch_size", C.c_int), ("msac_refinements", C.c_int),
                ("msac_threshold", C.c_double), ("seed", C.c_uint64), ("pyramid_levels", C.c_int),
                ("pyramid_scale_factor", C.c_double)]


class AttemptRecord(C.Structure):
    _fields_ = [("S", C.c_double), ("K", C.c_int), ("R", C.c_int), ("blend_mode", C.c_int), ("action", C.c_int),
                ("reason", C.c_int), ("next_S", C.c_double), ("next_K", C.c_int), ("next_R", C.c_int),
                ("next_blend_mode", C.c_int), ("previous_R", C.c_int), ("generation", C.c_uint64),
                ("homography", HomographyResult), ("t_s1", C.c_double), ("t_s2", C.c_double), ("t_s3", C.c_double),
                ("t_s4", C.c_double), ("t_s5", C.c_double), ("t_attempt", C.c_double),
                ("restart_from", C.c_char * 8), ("fail_stage", C.c_char * 8), ("execution_status", C.c_int),
                ("has_decision", C.c_int), ("has_homography", C.c_int), ("has_feedback", C.c_int)]


class StripRecord(C.Structure):
    _fields_ = [("status", C.c_int), ("fail_stage", C.c_char * 16), ("fail_pair", C.c_int),
                ("canvas_width", C.c_int), ("canvas_height", C.c_int), ("origin_x", C.c_int), ("origin_y", C.c_int),
                ("union_pixels", C.c_uint64), ("overlap_pixels", C.c_uint64),
                ("t_pairs", C.c_double), ("t_geometry", C.c_double), ("t_render", C.c_double),
                ("t_total", C.c_double)]


class RunRecord(C.Structure):
    _fields_ = [("S", C.c_double), ("K", C.c_int), ("R", C.c_int), ("blend_mode", C.c_int),
                ("t_s1", C.c_double), ("t_s2", C.c_double), ("t_s3", C.c_double), ("t_s4", C.c_double),
                ("t_s5", C.c_double), ("t_s6", C.c_double), ("t_controller", C.c_double), ("t_e2e", C.c_double),
                ("t_probe", C.c_double), ("t_stream_wait", C.c_double), ("t_gpu_span", C.c_double),
                ("probe_gradient_a", C.c_double), ("probe_gradient_b", C.c_double),
                ("attempt_count", C.c_int), ("num_kp_a", C.c_int), ("num_kp_b", C.c_int), ("num_matches", C.c_int),
                ("num_inliers", C.c_int), ("inlier_ratio", C.c_double), ("inlier_coverage", C.c_double),
                ("reproj_residual", C.c_double), ("H", C.c_double * 9), ("canvas_pixels", C.c_uint64),
                ("overlap_pixels", C.c_uint64), ("canvas_width", C.c_int), ("canvas_height", C.c_int),
                ("num_levels", C.c_int), ("cache_hit", C.c_int), ("status", C.c_char * 32),
                ("fail_stage", C.c_char * 16)]


class PairSpec(C.Structure):
    _fields_ = [("a", Image), ("b", Image), ("S", C.c_double), ("K", C.c_int), ("R", C.c_int),
                ("blend_mode", C.c_int), ("fast_threshold", C.c_int), ("harris_k", C.c_float),
                ("ratio_threshold", C.c_double), ("cross_check", C.c_int), ("msac_batch_size", C.c_int),
                ("msac_refinements", C.c_int), ("msac_threshold", C.c_double), ("seed", C.c_uint64)]


class PairOutcome(C.Structure):
    _fields_ = [("status", C.c_int), ("fail_stage", C.c_char * 16), ("num_kp_a", C.c_int), ("num_kp_b", C.c_int),
                ("scale_a", ScaleInfo), ("scale_b", ScaleInfo), ("homography", HomographyResult),
                ("geometry", Geometry), ("statistics", RenderStatistics), ("canvas", Image), ("mask", Image),
                ("keypoints_a", Keypoints), ("keypoints_b", Keypoints), ("descriptors_a", Descriptors),
                ("descriptors_b", Descriptors), ("matches", Matches), ("inlier_mask", C.c_void_p)]


class BatchTimings(C.Structure):
    _fields_ = [("feature_ms", C.c_double), ("render_ms", C.c_double), ("host_ms", C.c_double),
                ("total_ms", C.c_double)]


class KernelProbe(C.Structure):
    _fields_ = [("kernel", C.c_char * 32), ("level", C.c_int), ("pad", C.c_int), ("launches", C.c_int64),
                ("total_ms", C.c_double), ("algorithmic_bytes", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("s4_segment", C.c_int), ("workgroups_per_cu", C.c_int), ("graphs", C.c_int),
                ("stage_workgroups_per_cu", C.c_int * 6), ("s1_tile_rows", C.c_int)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_batch", C.c_int), ("max_live", C.c_int), ("adaptive", C.c_int),
                ("use_graphs", C.c_int), ("canvas_factor", C.c_double)]


class Request(C.Structure):
    _fields_ = [("id", C.c_uint64), ("a", Image), ("b", Image), ("arrival_ms", C.c_double), ("seed", C.c_uint64),
                ("fixed", PairSpec), ("keep_output", C.c_int)]


class Completion(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_char * 32), ("fail_stage", C.c_char * 16),
                ("arrival_ms", C.c_double), ("admitted_ms", C.c_double), ("done_ms", C.c_double),
                ("attempts", C.c_int), ("S", C.c_double), ("K", C.c_int), ("R", C.c_int), ("blend_mode", C.c_int),
                ("num_kp_a", C.c_int), ("num_kp_b", C.c_int), ("homography", HomographyResult),
                ("geometry", Geometry), ("statistics", RenderStatistics), ("canvas", Image), ("mask", Image),
                ("reserved_bytes", C.c_uint64)]


class ServeStats(C.Structure):
    _fields_ = [("completed", C.c_int), ("failed", C.c_int), ("makespan_ms", C.c_double), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("feature_launches", C.c_int),
                ("render_launches", C.c_int), ("mean_feature_batch", C.c_double),
                ("mean_render_batch", C.c_double), ("peak_pool_bytes", C.c_uint64),
                ("feature_busy_ms", C.c_double), ("render_busy_ms", C.c_double)]


P = C.POINTER
vp = C.c_void_p
sz = C.c_size_t

# name -> (restype, argtypes)
SIGNATURES = {
    "cs_version": (C.c_char_p, []),
    "cs_status_name": (C.c_char_p, [C.c_int]),
    "cs_make_scale_info": (C.c_int, [C.c_int, C.c_int, C.c_double, P(ScaleInfo)]),
    "cs_s1_workspace_bytes": (C.c_int, [C.c_int, C.c_int, C.c_double, P(sz)]),
    "cs_s1_preprocess": (C.c_int, [P(Image), P(Image), C.c_double, P(Image), P(Image), P(ScaleInfo), P(ScaleInfo),
                                   vp]),
    "cs_s1_preprocess_single": (C.c_int, [P(Image), C.c_double, P(Image), P(ScaleInfo), vp]),
    "cs_s1_gradient_workspace_bytes": (C.c_int, [P(sz)]),
    "cs_s1_compute_gradient": (C.c_int, [P(Image), P(Image), vp, vp, vp, vp]),
    "cs_s2_workspace_bytes": (C.c_int, [C.c_int, C.c_int, P(S2Config), P(sz)]),
    "cs_s2_detect": (C.c_int, [P(Image), P(S2Config), P(Keypoints), vp, sz, vp]),
    "cs_s3_workspace_bytes": (C.c_int, [C.c_int, C.c_int, P(sz)]),
    "cs_s3_describe": (C.c_int, [P(Image), P(Keypoints), P(Descriptors), vp, sz, vp]),
    "cs_s4_workspace_bytes": (C.c_int, [C.c_int, C.c_int, P(S4Config), P(sz)]),
    "cs_s4_knn_top2": (C.c_int, [P(Keypoints), P(Descriptors), P(Keypoints), P(Descriptors), vp, vp]),
    "cs_s4_match": (C.c_int, [P(Keypoints), P(Descriptors), P(Keypoints), P(Descriptors), P(S4Config), P(Matches),
                              vp, sz, vp]),
    "cs_s4_forward_neighbors": (C.c_int, [C.c_int, C.c_int, P(S4Config), vp, sz, P(vp)]),
    "cs_s4_match_noalloc": (C.c_int, [P(Keypoints), P(Descriptors), P(Keypoints), P(Descriptors), P(S4Config),
                                      P(Matches), vp]),
    "cs_s4_set_segment": (C.c_int, [C.c_int]),
    "cs_s5_workspace_bytes": (C.c_int, [P(S5Config), P(sz)]),
    "cs_s5_estimate": (C.c_int, [P(Keypoints), P(Keypoints), P(Matches), P(ScaleInfo), P(ScaleInfo), P(S5Config),
                                 vp, vp, vp, sz, vp]),
    "cs_s6_prepare_geometry": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, P(C.c_double), P(GeometryPolicy),
                                         P(Geometry)]),
    "cs_s6_effective_reductions": (C.c_int, [P(Geometry), C.c_int]),
    "cs_s6_workspace_bytes": (C.c_int, [P(Geometry), C.c_int, P(sz)]),
    "cs_s6_render": (C.c_int, [P(Image), P(Image), P(Geometry), C.c_int, P(Image), P(Image), vp, sz, vp,
                               P(RenderStatistics)]),
    "cs_s6_prepare_chain": (C.c_int, [C.c_int, P(C.c_int), P(C.c_int), P(C.c_double), P(GeometryPolicy),
                                      P(Geometry)]),
    "cs_s6_chain_workspace_bytes": (C.c_int, [C.c_int, P(Geometry), C.c_int, P(sz)]),
    "cs_s6_render_chain": (C.c_int, [C.c_int, P(Image), P(Geometry), C.c_int, P(Image), P(Image), vp, sz, vp,
                                     P(RenderStatistics)]),
    "cs_pyramid_level_counts": (C.c_int, [P(PyramidConfig), P(C.c_int)]),
    "cs_pyramid_workspace_bytes": (C.c_int, [C.c_int, C.c_int, C.c_double, P(PyramidConfig), P(sz)]),
    "cs_pyramid_detect_describe": (C.c_int, [P(Image), C.c_double, P(PyramidConfig), P(Keypoints), P(Descriptors),
                                             vp, P(ScaleInfo), vp, sz, vp]),
    "cs_extract_scene": (C.c_int, [P(Image), P(Image), vp, vp, vp, C.c_int]),
    "cs_extract_gate": (C.c_int, [P(Image), P(Image), P(Keypoints), P(Keypoints), P(Matches), vp,
                                  P(HomographyResult), P(ScaleInfo), P(ScaleInfo), vp, vp, vp, vp, C.c_uint]),
    "cs_strip_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P(vp)]),
    "cs_strip_destroy": (None, [vp]),
    "cs_strip_run": (C.c_int, [vp, C.c_int, P(Image), P(HarnessConfig), P(StripRecord), P(HomographyResult),
                               P(C.c_int)]),
    "cs_strip_outputs": (C.c_int, [vp, P(Image), P(Image), P(Geometry), C.c_int]),
    "cs_strip_stream": (vp, [vp]),
    "cs_runner_create": (vp, []),
    "cs_runner_destroy": (None, [vp]),
    "cs_runner_initialize": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_uint64]),
    "cs_runner_query_memory": (C.c_int, [C.c_int] * 7 + [P(C.c_uint64)] * 3),
    "cs_runner_upload_pair": (C.c_int, [vp, vp, C.c_int, C.c_int, vp, C.c_int, C.c_int]),
    "cs_runner_use_device_pair": (C.c_int, [vp, P(Image), P(Image)]),
    "cs_runner_execute_run": (C.c_int, [vp, P(HarnessConfig), P(RunRecord)]),
    "cs_runner_execute_adaptive_run": (C.c_int, [vp, vp, C.c_uint64, P(RunRecord)]),
    "cs_runner_set_flags": (C.c_int, [vp, C.c_int, C.c_int]),
    "cs_runner_outputs": (C.c_int, [vp, P(Image), P(Image), P(Keypoints), P(Keypoints), P(Descriptors),
                                    P(Descriptors), P(Matches), P(vp), P(HomographyResult)]),
    "cs_runner_stream": (vp, [vp]),
    "cs_runner_set_collect_features": (C.c_int, [vp, C.c_int]),
    "cs_runner_features": (C.c_int, [vp, vp, vp, P(C.c_int), P(C.c_double)]),
    "cs_runner_attempt_trace": (C.c_int, [vp, P(AttemptRecord), C.c_int, P(C.c_int)]),
    "cs_controller_create_legacy": (vp, [C.c_int]),
    "cs_controller_destroy": (None, [vp]),
    "cs_plan_retry": (C.c_int, [C.c_double, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                P(C.c_int)]),
    "cs_controller_walk": (C.c_int, [vp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                                     P(HomographyResult), P(C.c_int), C.c_int, P(C.c_double)]),
    "cs_controller_create_policy": (vp, [C.c_char_p, C.c_char_p, sz]),
    "cs_controller_observations": (C.c_int, [vp, P(C.c_int), P(C.c_int), P(C.c_uint), P(C.c_int)]),
    "cs_controller_walk_features": (C.c_int, [vp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, vp,
                                              P(HomographyResult), P(C.c_int), vp, P(C.c_int), C.c_int,
                                              P(C.c_double)]),
    "cs_synth_pair": (C.c_int, [C.c_uint64, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, vp, vp,
                                P(C.c_double)]),
    "cs_synth_strip": (C.c_int, [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, vp, P(C.c_double)]),
    "cs_batch_create": (C.c_int, [C.c_int, C.c_uint64, C.c_int, P(vp)]),
    "cs_batch_destroy": (None, [vp]),
    "cs_batch_run": (C.c_int, [vp, P(PairSpec), C.c_int, P(PairOutcome), P(BatchTimings)]),
    "cs_batch_submit": (C.c_int, [vp, P(PairSpec), C.c_int, P(PairOutcome), C.c_int, P(BatchTimings)]),
    "cs_batch_finish": (C.c_int, [vp, P(PairOutcome), C.c_int]),
    "cs_batch_autotune": (C.c_int, [vp, P(PairSpec), C.c_int, C.c_char_p, P(Tuning)]),
    "cs_batch_apply": (C.c_int, [vp, P(Tuning)]),
    "cs_batch_pool": (C.c_int, [vp, P(C.c_uint64), P(C.c_uint64), P(C.c_uint64)]),
    "cs_batch_set_probe": (C.c_int, [vp, C.c_int]),
    "cs_batch_probe_read": (C.c_int, [vp, P(KernelProbe), C.c_int, P(C.c_int)]),
    "cs_batch_pair_bytes": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, P(C.c_uint64)]),
    "cs_batch_render_bytes": (C.c_int, [C.c_int, C.c_int, C.c_int, P(C.c_uint64)]),
    "cs_server_create": (C.c_int, [P(ServerConfig), P(vp)]),
    "cs_server_create_policy": (C.c_int, [P(ServerConfig), C.c_char_p, P(vp)]),
    "cs_server_destroy": (None, [vp]),
    "cs_server_run_trace": (C.c_int, [vp, P(Request), C.c_int, P(Completion), P(ServeStats)]),
    "cs_server_release": (C.c_int, [vp, C.c_uint64]),
}

# Controller observation sizes (src/controller/features.hpp:10-13).
SCENE_FEATURES = 32
GATE_FEATURES = 38
FEATURE_DEVICE_DOUBLES = 136

_lib = None


def load() -> C.CDLL:
    """Loads libcsstitch.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python crossstage-stitch_amd/build.py`")
        lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | os.RTLD_GLOBAL)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib
