"""The plain per-pair runner (S1-S6 chained on one stream, two host syncs) on
synthetic pairs, checked stage by stage against the oracle chain: bit-exact
keypoints, descriptors, matches, identical RANSAC result and inlier mask,
panorama within 1 LSB with the union mask exact."""
import numpy as np
import pytest

import oracle_lib as O
import pipeline_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed,w,h,S,K,R,blend,cross", [
    (20261022, 512, 512, 1.0, 1000, 2000, "multiband_5", False),
    (20261023, 640, 480, 0.5, 800, 512, "feather", False),
    (20261024, 1280, 720, 0.75, 3000, 1024, "multiband_3", True),
    (20261025, 1920, 1080, 1.0, 5000, 4096, "multiband_6", True),
])
def test_plain_runner_matches_oracle(seed, w, h, S, K, R, blend, cross):
    from cs_stitch import pipeline as P
    a, b, H = P.synth_pair(seed, w, h, overlap=(0.3, 0.5) if w > 1000 else (0.6, 0.8))
    runner = P.PipelineRunner(w, h, max_k=K, max_r=R, canvas_width=4 * w, canvas_height=2 * h,
                              workspace_mode=P.F.MULTIBAND7)
    rec, ref = pipeline_oracle.check_runner(runner, a, b, S, K, R, blend, cross)
    assert ref["result"].status == 0
    # The estimate recovers the generator's homography: B pixels that land
    # inside A reproject within 1.5 px of the true mapping (median).
    Hs = np.array(rec.H[:]).reshape(3, 3)
    yy, xx = np.mgrid[0:h:16, 0:w:16]
    pts = np.stack([xx.ravel(), yy.ravel(), np.ones(xx.size)])
    p, q = Hs @ pts, H @ pts
    p, q = p[:2] / p[2], q[:2] / q[2]
    inside = (q[0] >= 0) & (q[0] < w) & (q[1] >= 0) & (q[1] < h)
    err = np.hypot(*(p - q))[inside]
    assert inside.sum() > 50 and np.median(err) < 1.5


def test_adaptive_runner_legacy_policy():
    """The adaptive path (probe -> rung -> attempts with S5 reuse -> S6) is
    deterministic and walks the legacy-v1 ladder. Every attempt's S5 result
    (status, inlier count, H bit for bit) equals the oracle's for the
    attempt's knobs and reuse plan (pipeline_oracle.replay_attempts), and an
    accepted pair's panorama is the oracle's render of the accepted H within
    1 LSB, union mask exact. Which terminal status a pair reaches follows
    from those results through the controller, whose decisions are pinned
    against the reference controller on the CPU (tests/test_controller.py)."""
    from cs_stitch import pipeline as P
    ladder = [(.25, 500, 128), (.5, 1000, 128), (.5, 1000, 256), (.75, 2000, 512), (1., 2000, 1024)]
    for seed, ov in ((7, (0.6, 0.8)), (8, (0.85, 0.95))):
        a, b, _ = P.synth_pair(seed, 640, 480, overlap=ov)
        runner = P.PipelineRunner(640, 480, max_k=4000, max_r=2048, canvas_width=2048, canvas_height=1024)
        runner.upload_pair(a, b)
        ctl = P.Controller()
        rec = runner.execute_adaptive_run(ctl, seed=0)
        status = rec.status.decode()
        print(seed, status, rec.attempt_count, rec.num_matches, rec.num_inliers, rec.inlier_ratio,
              rec.inlier_coverage, rec.reproj_residual, rec.probe_gradient_a, rec.probe_gradient_b)
        assert status in ("OK", "BUDGET_EXHAUSTED", "REJECTED"), (status, rec.fail_stage)
        assert 1 <= rec.attempt_count <= len(ladder)
        assert (rec.S, rec.K, rec.R) in ladder
        if status == "OK":
            assert rec.canvas_pixels > 0
        # The attempt trace walks the same ladder: each attempt's knobs are
        # the previous decision's next knobs, every attempt has an S5 result
        # and a decision, only the last one ends the run, and the first
        # attempt starts from scratch.
        trace = runner.attempt_trace()
        assert len(trace) == rec.attempt_count
        for i, at in enumerate(trace):
            assert (at.S, at.K, at.R) in ladder and at.has_homography and at.has_decision
            assert at.t_s5 > 0 and at.t_attempt >= at.t_s5
            if i + 1 < len(trace):
                assert at.action == 1 and (at.next_S, at.next_K, at.next_R) == (trace[i + 1].S, trace[i + 1].K,
                                                                                 trace[i + 1].R)
        assert trace[0].restart_from == b"none" and trace[0].previous_R == 0
        assert trace[-1].action == (0 if status == "OK" else 2)
        assert (trace[-1].S, trace[-1].K, trace[-1].R) == (rec.S, rec.K, rec.R)
        if status == "OK":
            assert trace[-1].homography.num_inliers == rec.num_inliers
        # Attempt by attempt against the oracle.
        for i, (t, want) in enumerate(zip(trace, pipeline_oracle.replay_attempts(a, b, trace))):
            got = t.homography
            assert (got.status, got.num_matches, got.num_inliers) == (want.status, want.num_matches,
                                                                      want.num_inliers), (seed, i)
            if want.status == 0:
                assert list(got.H) == list(want.H), (seed, i)
        if status == "OK":
            st_, geo = O.prepare_geometry(640, 480, 640, 480, np.array(rec.H[:]))
            assert st_ == 0
            pano, pmask, _ = O.s6_render(a, b, geo, trace[-1].blend_mode)
            out = runner.outputs()
            got = P.download_image(out["canvas"], 3)
            assert np.array_equal(P.download_image(out["mask"], 1), pmask)
            assert np.abs(got.astype(int) - pano.astype(int)).max() <= 1
        rec2 = runner.execute_adaptive_run(ctl, seed=0)
        assert rec2.status == rec.status and rec2.attempt_count == rec.attempt_count
        assert list(rec2.H) == list(rec.H) or status != "OK"
        runner.close()


def test_large_stress_adjacent_pair_matches_oracle():
    """One adjacent pair of the large-scale strip: 3840x2160 frames at 30 %
    overlap, 16000 keypoints per image (16000 x 16000 Hamming), RANSAC 8192,
    7-band blend; every stage against the oracle chain."""
    from cs_stitch import pipeline as P
    w, h = 3840, 2160
    # A seed whose pair canvas (6046 x 2289) fits the 16.7 MP canvas budget;
    # the whole 8-frame strip does not (it is not rendered here).
    a, b, _ = P.synth_pair(20261204, w, h, overlap=(0.3, 0.3))
    runner = P.PipelineRunner(w, h, max_k=16000, max_r=8192, canvas_width=6400, canvas_height=2600,
                              workspace_mode=P.F.MULTIBAND7)
    rec, ref = pipeline_oracle.check_runner(runner, a, b, 1.0, 16000, 8192, "multiband_7", False)
    print("stress pair", rec.status, rec.num_kp_a, rec.num_kp_b, rec.num_matches, rec.num_inliers,
          rec.canvas_width, rec.canvas_height)
    assert ref["result"].status == 0 and rec.status == b"OK"
    runner.close()
