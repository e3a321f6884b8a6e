"""Strip stitching (cs_strip_run) on the GPU: every adjacent pair's S1-S5
result equals the plain runner's on the same two frames bit for bit, the
recovered homographies land within two pixels of the true ones over the
overlap, the chain geometry equals the oracle's for the same homographies,
and the canvas is within 1 LSB of the oracle's full-canvas chain render with the mask and the
union/overlap counts exact."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

INT_FIELDS = ["width", "height", "origin_x", "origin_y", "width_a", "height_a", "width_b", "height_b"]


@pytest.fixture(scope="module")
def T():
    import torch
    from cs_stitch import pipeline, stages
    assert torch.cuda.is_available()
    return torch, stages, pipeline


def dev_rgb(torch, st, arr, pad=64):
    t = st.pitched(arr.shape[0], arr.shape[1], 3, pad)
    t.copy_(torch.from_numpy(np.ascontiguousarray(arr)).cuda())
    return t


def overlap_error(h_est, h_true, w, h):
    """Median distance between the estimated and true images of a 16-pixel
    grid of frame k + 1 over the part that lands in frame k."""
    errs = []
    for y in range(0, h, 16):
        for x in range(0, w, 16):
            q = h_true @ np.array([x, y, 1.0])
            u, v = q[:2] / q[2]
            if 0 <= u < w and 0 <= v < h:
                p = h_est @ np.array([x, y, 1.0])
                errs.append(float(np.hypot(*(p[:2] / p[2] - (u, v)))))
    return float(np.median(errs))


@pytest.mark.parametrize("blend", ["multiband_5", "feather"])
def test_strip_matches_runner_and_oracle(T, blend):
    torch, st, pl = T
    n, w, h = 4, 640, 480
    frames, truth = pl.synth_strip(50, n, w, h, 0.45)
    cfg = pl.harness_config(S=1.0, K=2000, R=1024, blend=blend, cross_check=True)
    dev = [dev_rgb(torch, st, f) for f in frames]
    print(f"strip n={n} {w}x{h} {blend}", flush=True)
    strip = pl.StripStitcher(n, w, h, max_k=2000, max_r=1024)
    status, rec, pairs, counts = strip.run(dev, cfg)
    assert status == 0, (rec.fail_stage, rec.fail_pair, status)
    assert all(c > 50 for c in counts)
    runner = pl.PipelineRunner(w, h, max_k=2000, max_r=1024, canvas_width=2048, canvas_height=1024)
    for k in range(n - 1):
        runner.use_device_pair(dev[k], dev[k + 1])
        r = runner.execute_run(cfg)
        assert r.status.decode() == "OK", k
        got = pairs[k]
        assert np.array(got.H).tobytes() == np.array(r.H).tobytes(), k
        assert (got.num_matches, got.num_inliers) == (r.num_matches, r.num_inliers), k
        assert (counts[k], counts[k + 1]) == (r.num_kp_a, r.num_kp_b), k
        # The synthetic frames give about a hundred keypoints each, so the
        # estimate is good to a pixel or two over the overlap.
        err = overlap_error(np.array(got.H).reshape(3, 3), truth[k], w, h)
        assert err < 2.0, (k, err)
    hs = [np.array(p.H).reshape(3, 3) for p in pairs]
    canvas, mask, geoms = strip.outputs(n)
    ost, og = O.prepare_chain([(w, h)] * n, hs, max_pixels=1 << 28)
    assert ost == 0
    for k in range(n):
        assert [getattr(geoms[k], f) for f in INT_FIELDS] == [getattr(og[k], f) for f in INT_FIELDS], k
        assert np.array(geoms[k].inverse).tobytes() == np.array(og[k].inverse).tobytes(), k
    assert (rec.canvas_width, rec.canvas_height) == (og[0].width, og[0].height)
    ref_out, ref_mask, ref_stats = O.s6_render_chain(list(frames), og, O.MODES[blend])
    got_out = pl.download_image(canvas, 3)
    got_mask = pl.download_image(mask, 1)
    assert np.array_equal(got_mask, ref_mask)
    assert (rec.union_pixels, rec.overlap_pixels) == tuple(ref_stats)
    diff = np.abs(got_out.astype(int) - ref_out.astype(int))
    assert diff.max() <= 1 and (diff > 0).mean() < 1e-3, (int(diff.max()), float((diff > 0).mean()))
    strip.close()
    runner.close()


def test_strip_reports_a_failing_pair(T):
    torch, st, pl = T
    # Frame 2 is flat: its pairs have no keypoints, and S5 of pair 1 fails.
    frames, _ = pl.synth_strip(51, 3, 320, 240, 0.5)
    frames[2][:] = 128
    dev = [dev_rgb(torch, st, f) for f in frames]
    strip = pl.StripStitcher(3, 320, 240, max_k=1000, max_r=256)
    status, rec, pairs, counts = strip.run(dev, pl.harness_config(S=1.0, K=1000, R=256))
    assert status != 0 and rec.fail_stage.decode() == "s5" and rec.fail_pair == 1
    assert pairs[0].status == 0 and counts[2] == 0
    strip.close()
