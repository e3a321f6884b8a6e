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


def replay_attempts(a, b, trace, seed=0, fast=20, ratio=0.8, threshold=3.0, refinements=1):
    """The adaptive run's attempts (harness/pipeline_runner.cu:531-885) on the
    oracle, from the knobs and reuse plan the run recorded: an attempt that
    restarts (the first one, or restart_from S1/S2) runs S1-S4 at its S and
    K without cross-check, one that restarts at S5 keeps the previous
    matches; S5 then evaluates hypotheses [previous_R, R) from the state the
    earlier attempts left (a fresh state when previous_R is 0). Returns the
    oracle's S5 result per attempt and the matching/feature state of the
    last one."""
    results, feats, state = [], None, None
    for t in trace:
        restart = t.restart_from.decode()
        if feats is None or restart in ("none", "S1", "S2"):
            ga, gb = O.s1_preprocess(a, t.S), O.s1_preprocess(b, t.S)
            ka, kb = O.s2_detect(ga, fast, t.K), O.s2_detect(gb, fast, t.K)
            da, db = O.s3_describe(ga, ka), O.s3_describe(gb, kb)
            ia, ib, _ = O.s4_match(da, db, ratio, False)
            _, sa = O.scale_info(a.shape[1], a.shape[0], t.S)
            _, sb = O.scale_info(b.shape[1], b.shape[0], t.S)
            feats = (ka, kb, ia, ib, sa, sb)
        ka, kb, ia, ib, sa, sb = feats
        if t.previous_R == 0:
            state = None
        res, _, state = O.s5_estimate_kp(ka[:, 0], ka[:, 1], kb[:, 0], kb[:, 1], ia, ib, sa, sb, t.R, refinements,
                                         threshold, seed, previous_hypotheses=t.previous_R, state=state)
        results.append(res)
    return results


def check_outputs(d, ref, what=""):
    """Host copies of one pair's device outputs (keys as tests/test_gpu_batch
    download_outcome: kp_a/kp_b (n, 3) float32, desc_a/desc_b bytes, nm,
    idx_a/idx_b/dist bytes, H, h_status, inliers, mask_inl, canvas, mask)
    against the oracle chain: keypoints, descriptors and matches bit-exact,
    the S5 status, inlier count, H and inlier mask identical, and - when S5
    and the geometry succeed - the panorama within 1 LSB with the union mask
    exact."""
    assert np.array_equal(d["kp_a"], ref["kp_a"]), (what, "kp_a")
    assert np.array_equal(d["kp_b"], ref["kp_b"]), (what, "kp_b")
    assert np.array_equal(np.asarray(d["desc_a"]).reshape(-1, 32), ref["desc_a"]), (what, "desc_a")
    assert np.array_equal(np.asarray(d["desc_b"]).reshape(-1, 32), ref["desc_b"]), (what, "desc_b")
    nm = int(d["nm"])
    assert nm == len(ref["idx_a"]) == ref["result"].num_matches, (what, nm, len(ref["idx_a"]))
    assert np.array_equal(np.asarray(d["idx_a"]).view(np.int32), ref["idx_a"]), (what, "idx_a")
    assert np.array_equal(np.asarray(d["idx_b"]).view(np.int32), ref["idx_b"]), (what, "idx_b")
    assert np.array_equal(np.asarray(d["dist"]).view(np.float32), ref["dist"]), (what, "distance")
    res = ref["result"]
    assert d["h_status"] == res.status and d["inliers"] == res.num_inliers, (what, d["h_status"], res.status)
    if res.status == 0:
        assert np.array_equal(np.asarray(d["H"]), np.array(res.H[:])), (what, "H")
    assert np.array_equal(np.asarray(d["mask_inl"]), ref["mask"][:nm]), (what, "inlier mask")
    if res.status != 0 or ref.get("geometry_status", 1) != 0:
        return
    pano, pmask = d["canvas"], d["mask"]
    assert pano.shape == ref["panorama"].shape, (what, pano.shape, ref["panorama"].shape)
    assert np.array_equal(pmask, ref["pano_mask"]), (what, "union mask")
    assert np.abs(pano.astype(int) - ref["panorama"].astype(int)).max() <= 1, (what, "panorama")


def check_runner(runner, a, b, S, K, R, blend, cross):
    """Uploads the pair, runs the plain runner once and checks every output
    against the oracle chain: bit-exact keypoints, descriptors and matches,
    identical S5 result and inlier mask, panorama within 1 LSB with the union
    mask exact. Returns (run record, oracle outputs)."""
    from cs_stitch import pipeline as P
    runner.upload_pair(a, b)
    rec = runner.execute_run(P.harness_config(S=S, K=K, R=R, blend=blend, cross_check=cross))
    ref = run(a, b, S, K, R, blend, cross)
    out = runner.outputs()
    na, nb = rec.num_kp_a, rec.num_kp_b
    assert na == ref["kp_a"].shape[0] and nb == ref["kp_b"].shape[0]
    for key, kp, n in (("kp_a", out["ka"], na), ("kp_b", out["kb"], nb)):
        xs = P.download(kp.x, 4 * n).view(np.float32)
        ys = P.download(kp.y, 4 * n).view(np.float32)
        rs = P.download(kp.response, 4 * n).view(np.float32)
        assert np.array_equal(np.stack([xs, ys, rs], 1), ref[key]), key
    assert np.array_equal(P.download(out["da"].data, 32 * na).reshape(-1, 32), ref["desc_a"])
    assert np.array_equal(P.download(out["db"].data, 32 * nb).reshape(-1, 32), ref["desc_b"])
    res = ref["result"]
    assert rec.num_matches == res.num_matches == len(ref["idx_a"])
    m = out["matches"]
    nm = rec.num_matches
    assert np.array_equal(P.download(m.idx_a, 4 * nm).view(np.int32), ref["idx_a"])
    assert np.array_equal(P.download(m.idx_b, 4 * nm).view(np.int32), ref["idx_b"])
    assert np.array_equal(P.download(m.distance, 4 * nm).view(np.float32), ref["dist"])
    h = out["h"]
    assert h.status == res.status and h.num_inliers == res.num_inliers
    assert np.array_equal(np.array(h.H[:]), np.array(res.H[:])) or res.status != 0
    assert np.array_equal(P.download(out["inlier_mask"], nm), ref["mask"][:nm])
    if res.status != 0:
        return rec, ref
    assert rec.status.decode() == "OK", rec.status
    pano = P.download_image(out["canvas"], 3)
    pmask = P.download_image(out["mask"], 1)
    assert pano.shape == ref["panorama"].shape
    assert np.array_equal(pmask, ref["pano_mask"])
    assert np.abs(pano.astype(int) - ref["panorama"].astype(int)).max() <= 1
    assert rec.overlap_pixels == ref["stats"][1]
    return rec, ref
