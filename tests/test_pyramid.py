"""Multi-level detection (an S1 pyramid over S2/S3), a build-defined
extension: the reference detects on one level only, so nothing of the
reference pins levels above 0 ("parity unpinned" against the reference).

The oracle is the composition of the pinned single-level oracle stages: S1
at scale S / f^l, S2 with the level's share of K, S3, and the map back to
level-0 work coordinates. The level split is checked on closed forms
(f = 2 gives exact geometric shares). GPU results must equal the
composition bit for bit.
"""
import numpy as np
import pytest

import oracle_lib as O


def split(levels, f, k):
    """Keypoints per level: rounded geometric shares, remainder last."""
    ratio = 1.0 / f
    per = k * (1.0 - ratio) / (1.0 - ratio ** levels)
    out, total = [], 0
    for _ in range(levels - 1):
        c = int(np.rint(per))
        out.append(c)
        total += c
        per *= ratio
    out.append(max(k - total, 0))
    return out


def test_level_split_closed_forms():
    from cs_stitch import stages as st
    assert st.pyramid_level_counts(2, 2.0, 3000) == [2000, 1000]
    assert st.pyramid_level_counts(3, 2.0, 7) == [4, 2, 1]
    assert st.pyramid_level_counts(1, 1.2, 5000) == [5000]
    for levels, f, k in ((8, 1.2, 5000), (4, 1.5, 1000), (16, 1.05, 37), (8, 1.2, 0)):
        c = st.pyramid_level_counts(levels, f, k)
        assert c == split(levels, f, k) and sum(c) == k
        assert all(a >= b for a, b in zip(c[:-2], c[1:-1]))


def oracle_pyramid(rgb, S, levels, f, thr, k, harris_k=0.04):
    h, w, _ = rgb.shape
    _, s0 = O.scale_info(w, h, S)
    xs, ys, rs, ds, lv = [], [], [], [], []
    for l, n in enumerate(split(levels, f, k)):
        scale = S / f ** l
        st, sl = O.scale_info(w, h, scale)
        if n == 0 or st != 0 or sl.work_w < 64 or sl.work_h < 64:
            continue
        g = O.s1_preprocess(rgb, scale)
        kp = O.s2_detect(g, thr, n, harris_k)
        d = O.s3_describe(g, kp)
        x = kp[:, 0].astype(np.float64)
        y = kp[:, 1].astype(np.float64)
        xs.append(((x + 0.5) / sl.sx * s0.sx - 0.5).astype(np.float32))
        ys.append(((y + 0.5) / sl.sy * s0.sy - 0.5).astype(np.float32))
        rs.append(kp[:, 2])
        ds.append(d)
        lv.append(np.full(len(kp), l, np.uint8))
    cat = lambda a, dt: np.concatenate(a) if a else np.zeros(0, dt)
    return (cat(xs, np.float32), cat(ys, np.float32), cat(rs, np.float32),
            np.concatenate(ds) if ds else np.zeros((0, 32), np.uint8), cat(lv, np.uint8))


@pytest.mark.gpu
@pytest.mark.parametrize("w,h,S,levels,f,k", [(640, 480, 1.0, 4, 1.2, 2000), (1280, 720, 0.5, 8, 1.2, 3000),
                                              (800, 600, 0.75, 3, 2.0, 1000), (640, 480, 1.0, 1, 1.2, 1500)])
def test_pyramid_matches_composed_oracle(w, h, S, levels, f, k):
    import torch
    from cs_stitch import pipeline as P
    from cs_stitch import stages as st
    a, _, _ = P.synth_pair(6100 + w, w, h, overlap=(0.5, 0.7))
    t = st.pitched(h, w, 3)
    t.copy_(torch.from_numpy(a).cuda())
    kp, desc, lv, si = st.pyramid_detect_describe(t, S, levels, f, 20, k)
    torch.cuda.synchronize()
    n, x, y, r = kp.host()
    ox, oy, orr, od, ol = oracle_pyramid(a, S, levels, f, 20, k)
    print(w, h, S, levels, f, k, "keypoints", n, "per level", np.bincount(ol, minlength=levels).tolist())
    assert n == len(ox) and n > 0
    assert np.array_equal(x.numpy(), ox) and np.array_equal(y.numpy(), oy) and np.array_equal(r.numpy(), orr)
    assert np.array_equal(desc[:n].cpu().numpy(), od)
    assert np.array_equal(lv[:n].cpu().numpy(), ol)
    _, s0 = O.scale_info(w, h, S)
    assert (si.work_w, si.work_h) == (s0.work_w, s0.work_h)
    # Every keypoint lies inside the level-0 image.
    assert x.min() >= 0 and y.min() >= 0 and x.max() < s0.work_w and y.max() < s0.work_h


@pytest.mark.gpu
def test_runner_with_pyramid_matches_oracle_chain():
    """The plain runner with pyramid_levels = 4: keypoints, descriptors and
    matches of the pyramid, then S5, equal the oracle chain run on the
    composed pyramid."""
    from cs_stitch import pipeline as P
    w, h, S, K, R, levels = 960, 720, 1.0, 2000, 512, 4
    a, b, _ = P.synth_pair(6200, w, h, overlap=(0.5, 0.7))
    runner = P.PipelineRunner(w, h, max_k=4000, max_r=2048, canvas_width=4096, canvas_height=2048)
    runner.upload_pair(a, b)
    rec = runner.execute_run(P.harness_config(S=S, K=K, R=R, pyramid_levels=levels, pyramid_scale_factor=1.2))
    pa = oracle_pyramid(a, S, levels, 1.2, 20, K)
    pb = oracle_pyramid(b, S, levels, 1.2, 20, K)
    assert (rec.num_kp_a, rec.num_kp_b) == (len(pa[0]), len(pb[0]))
    ia, ib, dist = O.s4_match(pa[3], pb[3], 0.8, False)
    _, sa = O.scale_info(w, h, S)
    _, sb = O.scale_info(w, h, S)
    res, mask, _ = O.s5_estimate_kp(pa[0], pa[1], pb[0], pb[1], ia, ib, sa, sb, R, 1, 3.0, 0)
    out = runner.outputs()
    nm = rec.num_matches
    print("pyramid runner", rec.status, rec.num_kp_a, rec.num_kp_b, nm, rec.num_inliers)
    assert nm == len(ia) and rec.num_inliers == res.num_inliers
    assert np.array_equal(P.download(out["matches"].idx_a, 4 * nm).view(np.int32), ia)
    assert np.array_equal(P.download(out["matches"].idx_b, 4 * nm).view(np.int32), ib)
    assert np.array_equal(P.download(out["inlier_mask"], nm), mask[:nm])
    assert out["h"].status == res.status
    if res.status == 0:
        assert np.array_equal(np.array(out["h"].H[:]), np.array(res.H[:]))
        assert rec.status == b"OK"
    runner.close()


@pytest.mark.gpu
def test_reference_cpu_case_with_pyramid_end_to_end():
    """The reference's CPU-runnable case: 512x512, FAST 20, 1000 keypoints on
    an 8-level pyramid at scale 1.2, ratio 0.8, RANSAC 2000 at 3 px, 5-band
    blend. Every stage equals the oracle chain on the composed pyramid; the
    panorama is within 1 LSB with the mask exact."""
    from cs_stitch import pipeline as P
    w = h = 512
    a, b, _ = P.synth_pair(20261201, w, h, overlap=(0.6, 0.8))
    runner = P.PipelineRunner(w, h, max_k=4000, max_r=2048, canvas_width=2048, canvas_height=2048)
    runner.upload_pair(a, b)
    rec = runner.execute_run(P.harness_config(S=1.0, K=1000, R=2000, blend="multiband_5", pyramid_levels=8,
                                              pyramid_scale_factor=1.2))
    pa = oracle_pyramid(a, 1.0, 8, 1.2, 20, 1000)
    pb = oracle_pyramid(b, 1.0, 8, 1.2, 20, 1000)
    ia, ib, _ = O.s4_match(pa[3], pb[3], 0.8, False)
    _, s0 = O.scale_info(w, h, 1.0)
    res, mask, _ = O.s5_estimate_kp(pa[0], pa[1], pb[0], pb[1], ia, ib, s0, s0, 2000, 1, 3.0, 0)
    print("reference CPU case", rec.status, rec.num_kp_a, rec.num_kp_b, rec.num_matches, rec.num_inliers)
    assert (rec.num_kp_a, rec.num_kp_b, rec.num_matches) == (len(pa[0]), len(pb[0]), len(ia))
    assert res.status == 0 and rec.status == b"OK" and rec.num_inliers == res.num_inliers
    out = runner.outputs()
    assert np.array_equal(np.array(out["h"].H[:]), np.array(res.H[:]))
    st, geo = O.prepare_geometry(w, h, w, h, np.array(res.H[:]))
    assert st == 0
    pano, pmask, _ = O.s6_render(a, b, geo, O.MODES["multiband_5"])
    got = P.download_image(out["canvas"], 3)
    assert np.array_equal(P.download_image(out["mask"], 1), pmask)
    assert np.abs(got.astype(int) - pano.astype(int)).max() <= 1
    runner.close()
