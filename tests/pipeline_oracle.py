"""Runs the whole per-pair path (S1-S6) on the CPU oracle, the way the plain
runner chains the stages (harness/pipeline_runner.cu:308-529)."""
import numpy as np

import oracle_lib as O


def run(a, b, S, K, R, blend, cross_check, fast=20, ratio=0.8, refinements=1, threshold=3.0, seed=0):
    ga, gb = O.s1_preprocess(a, S), O.s1_preprocess(b, S)
    ka, kb = O.s2_detect(ga, fast, K), O.s2_detect(gb, fast, K)
    da, db = O.s3_describe(ga, ka), O.s3_describe(gb, kb)
    ia, ib, dist = O.s4_match(da, db, ratio, cross_check)
    _, sa = O.scale_info(a.shape[1], a.shape[0], S)
    _, sb = O.scale_info(b.shape[1], b.shape[0], S)
    res, mask, _ = O.s5_estimate_kp(ka[:, 0], ka[:, 1], kb[:, 0], kb[:, 1], ia, ib, sa, sb, R, refinements, threshold,
                                    seed)
    out = dict(gray_a=ga, gray_b=gb, kp_a=ka, kp_b=kb, desc_a=da, desc_b=db, idx_a=ia, idx_b=ib, dist=dist,
               result=res, mask=mask)
    if res.status != 0:
        return out
    st, geo = O.prepare_geometry(a.shape[1], a.shape[0], b.shape[1], b.shape[0], np.array(res.H[:]))
    out["geometry_status"] = st
    if st == 0:
        pano, pmask, stats = O.s6_render(a, b, geo, O.MODES[blend] if isinstance(blend, str) else blend)
        out.update(panorama=pano, pano_mask=pmask, stats=stats, geometry=geo)
    return out
