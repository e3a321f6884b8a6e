"""The batched and served forms at the sizes they are measured at.

Batched: one ragged batch shaped like the 64-pair mixed workload - four
3840x2160 pairs (one of whose canvases exceeds the 16,777,216-pixel budget
and must be rejected at S6, as harness/pipeline_runner.cu:464-470 rejects
it), 1920x1080, 1280x720 and 640x480 pairs, keypoint caps from 500 to 8000,
RANSAC 4096, 6-band blending and cross-checked matching. Three pairs (a 4K
pair at K 8000, a 1080p pair at K 8000 and a 720p pair at K 7000) are checked
stage by stage against the oracle chain (the reference's per-pair flow,
harness/pipeline_runner.cu:308-529); every pair equals the plain per-pair
runner bit for bit.

Served: an HBM budget that holds about two and a half requests' reservations
(the per-request byte plan of harness/pipeline_runner.cu:110-136), so
requests queue for memory and are admitted as earlier ones complete. The
trace mixes a 4K pair with low-texture, low-overlap and repetitive scenes;
every request ends exactly as it does alone.
"""
import concurrent.futures as cf

import numpy as np
import pytest
import torch

import pipeline_oracle
from test_gpu_batch import download_outcome, outcome_dict, same

pytestmark = pytest.mark.gpu

R, BLEND = 4096, "multiband_6"
# (seed, width, height, K, overlap range)
MIXED = [
    (31100, 3840, 2160, 8000, (0.3, 0.5)),
    (31101, 3840, 2160, 6000, (0.5, 0.8)),
    (31009, 3840, 2160, 3000, (0.2, 0.25)),  # canvas ~7500x2850 (21 MP): over the budget
    (31117, 3840, 2160, 7500, (0.4, 0.8)),
    (31104, 1920, 1080, 8000, (0.3, 0.5)),
    (31105, 1920, 1080, 5000, (0.2, 0.8)),
    (31106, 1920, 1080, 2500, (0.2, 0.8)),
    (31107, 1920, 1080, 700, (0.2, 0.8)),
    (31108, 1280, 720, 4000, (0.2, 0.8)),
    (31109, 1280, 720, 1200, (0.2, 0.8)),
    (31110, 1280, 720, 7000, (0.3, 0.6)),
    (31111, 1280, 720, 500, (0.2, 0.8)),
    (31112, 640, 480, 500, (0.2, 0.8)),
    (31113, 640, 480, 2000, (0.2, 0.8)),
    (31114, 640, 480, 8000, (0.2, 0.8)),
    (31115, 640, 480, 1000, (0.2, 0.8)),
]
OVERFLOW = 2
ORACLE_CHECKED = (0, 4, 10)
INSUFFICIENT_WORKSPACE = 2


def upload(a, b):
    from cs_stitch import stages as st
    h, w = a.shape[:2]
    ta, tb = st.pitched(h, w, 3), st.pitched(h, w, 3)
    ta.copy_(torch.from_numpy(a).cuda())
    tb.copy_(torch.from_numpy(b).cuda())
    return ta, tb


def test_mixed_batch_matches_oracle_and_plain_runner():
    from cs_stitch import batch as B
    from cs_stitch import pipeline as P
    host, dev = [], []
    for seed, w, h, k, ov in MIXED:
        a, b, _ = P.synth_pair(seed, w, h, overlap=ov)
        host.append((a, b))
        dev.append(upload(a, b))
    torch.cuda.synchronize()
    cfgs = [dict(S=1.0, K=k, R=R, blend=BLEND, cross_check=True, seed=0) for (_, _, _, k, _) in MIXED]
    # The oracle chains run on host threads (the C calls release the GIL)
    # while the GPU works.
    pool = cf.ThreadPoolExecutor(len(ORACLE_CHECKED))
    futures = {i: pool.submit(pipeline_oracle.run, *host[i], 1.0, cfgs[i]["K"], R, BLEND, True) for i in ORACLE_CHECKED}
    br = B.BatchRunner(max_pairs=len(MIXED), pool_bytes=32 << 30)
    specs = [B.pair_spec(ta, tb, **c) for (ta, tb), c in zip(dev, cfgs)]
    outs, t = br.run(specs)
    print("batch ms", t.feature_ms, t.render_ms, t.host_ms, t.total_ms, flush=True)
    got = [download_outcome(outcome_dict(o), c["K"]) for o, c in zip(outs, cfgs)]
    statuses = [o.status for o in outs]
    stages = [o.fail_stage.decode() for o in outs]
    print("statuses", list(zip(statuses, stages)), flush=True)
    assert statuses[OVERFLOW] == INSUFFICIENT_WORKSPACE and stages[OVERFLOW] == "s6"
    assert outs[OVERFLOW].homography.status == 0
    assert sum(s == 0 for s in statuses) >= 12
    for i in ORACLE_CHECKED:
        assert statuses[i] == 0, (i, statuses[i], stages[i])
    # Every pair through the plain per-pair runner, one runner per pair sized
    # to the pair (its canvas box is the batch's geometry, or the input size
    # when the batch rendered nothing).
    for i, ((ta, tb), c, o) in enumerate(zip(dev, cfgs, outs)):
        w, h = MIXED[i][1], MIXED[i][2]
        cw, ch = (o.geometry.width, o.geometry.height) if o.status == 0 else (w, h)
        runner = P.PipelineRunner(w, h, max_k=c["K"], max_r=R, canvas_width=cw, canvas_height=ch,
                                  workspace_mode=P.F.BLEND_NAMES[BLEND])
        runner.use_device_pair(ta, tb)
        rec = runner.execute_run(P.harness_config(S=1.0, K=c["K"], R=R, blend=BLEND, cross_check=True, seed=0))
        out = runner.outputs()
        plain = dict(na=rec.num_kp_a, nb=rec.num_kp_b, ka=out["ka"], kb=out["kb"], da=out["da"], db=out["db"],
                     matches=out["matches"], h=out["h"], inlier_mask=out["inlier_mask"],
                     canvas=out["canvas"] if rec.status == b"OK" else None, mask=out["mask"])
        dp = download_outcome(plain, c["K"])
        print(i, MIXED[i][1:4], rec.status, dp["nm"], dp["inliers"], flush=True)
        if i == OVERFLOW:
            assert rec.status == b"ERR_CANVAS_OVERFLOW" and rec.fail_stage == b"s6"
        else:
            assert (rec.status == b"OK") == (o.status == 0), (i, rec.status, o.status)
        same(dp, got[i], f"pair {i}")
        runner.close()
    br.close()
    for i in ORACLE_CHECKED:
        ref = futures[i].result()
        pipeline_oracle.check_outputs(got[i], ref, f"pair {i}")
        print("oracle ok", i, MIXED[i][1:4], ref["result"].num_matches, ref["result"].num_inliers, flush=True)
    pool.shutdown()
    # The oracle sees the same canvas overflow.
    import oracle_lib as O
    st, _ = O.prepare_geometry(3840, 2160, 3840, 2160, np.array(outs[OVERFLOW].homography.H[:]))
    assert st == INSUFFICIENT_WORKSPACE


# (seed, width, height, scene kind, overlap range): 0 textured, 1 low texture
# (blob density / 10), 2 repetitive (tiled pattern)
SERVED = [
    (32000, 3840, 2160, 0, (0.4, 0.6)),
    (32001, 1280, 720, 1, (0.4, 0.6)),
    (32002, 1920, 1080, 0, (0.1, 0.25)),
    (32003, 1280, 720, 2, (0.4, 0.6)),
    (32004, 1920, 1080, 0, (0.3, 0.6)),
    (32005, 640, 480, 1, (0.2, 0.5)),
    (32006, 1920, 1080, 2, (0.3, 0.6)),
]


def result_bits(c, with_canvas):
    """What a completion says about its request; the canvas only when the
    request kept it (a released canvas's memory may already serve another)."""
    from cs_stitch import pipeline as P
    d = dict(status=c.status, fail_stage=c.fail_stage, attempts=c.attempts, knobs=(c.S, c.K, c.R, c.blend_mode),
             H=np.array(c.homography.H[:]), h_status=c.homography.status, inliers=c.homography.num_inliers,
             matches=c.homography.num_matches, kp=(c.num_kp_a, c.num_kp_b))
    if c.status == b"OK":
        d["geometry"] = (c.geometry.width, c.geometry.height, c.geometry.origin_x, c.geometry.origin_y)
        d["stats"] = (c.statistics.union_pixels, c.statistics.overlap_pixels)
        if with_canvas:
            d["canvas"] = P.download_image(c.canvas, 3)
            d["mask"] = P.download_image(c.mask, 1)
    return d


def test_server_admission_waits_for_memory():
    from cs_stitch import batch as B
    from cs_stitch import pipeline as P
    dev = []
    for seed, w, h, kind, ov in SERVED:
        a, b, _ = P.synth_pair(seed, w, h, overlap=ov, kind=kind)
        dev.append(upload(a, b))
    torch.cuda.synchronize()
    n = len(SERVED)
    kept = n - 1  # admitted last, so keeping its canvas holds nobody back

    def requests(ids):
        return [B.request(i, *dev[i], 0.0, seed=5, keep_output=i == kept) for i in ids]

    alone, reserved = {}, {}
    srv = B.Server(hbm_budget=16 << 30, max_batch=8)
    for i in range(n):
        (c,), _ = srv.run_trace(requests([i]))
        alone[i] = result_bits(c, i == kept)
        reserved[i] = c.reserved_bytes
        srv.release(i)
    srv.close()
    print("alone", [(i, alone[i]["status"], alone[i]["attempts"], reserved[i] >> 20) for i in range(n)], flush=True)
    ordered = sorted(reserved.values())
    budget = max(int(2.5 * ordered[n // 2]), ordered[-1] + (64 << 20))
    assert sum(reserved.values()) > budget  # not everything fits at once
    srv = B.Server(hbm_budget=budget, max_batch=8)
    mix, st = srv.run_trace(requests(range(n)))
    print("mix", [(c.id, c.status, round(c.admitted_ms, 2), round(c.done_ms, 2)) for c in mix],
          "budget MiB", budget >> 20, "peak MiB", st.peak_pool_bytes >> 20, flush=True)
    assert len(mix) == n and sorted(c.id for c in mix) == list(range(n))
    by_id = {c.id: c for c in mix}
    # Memory: when any request was admitted, the reservations of the requests
    # then in flight (admitted, not yet done) fit the budget.
    for c in mix:
        t = c.admitted_ms
        live = [d for d in mix if d.admitted_ms <= t < d.done_ms or d.id == c.id]
        assert sum(d.reserved_bytes for d in live) <= budget, (c.id, [d.id for d in live])
    assert st.peak_pool_bytes <= budget
    # Admission waited for memory: some requests were admitted only after an
    # earlier request completed and freed its reservation.
    first_done = min(c.done_ms for c in mix)
    waited = [c for c in mix if c.admitted_ms >= first_done]
    assert waited, "every request was admitted at once"
    for c in waited:
        assert any(d.done_ms <= c.admitted_ms for d in mix if d.id != c.id)
        assert c.admitted_ms > c.arrival_ms
    # Each result is the one it has alone (the canvas of the kept request too).
    for i in range(n):
        same(alone[i], result_bits(by_id[i], i == kept), f"request {i}")
    srv.release(kept)
    srv.close()
