"""Batched and served forms.

Batched: a ragged batch (mixed resolutions, keypoint caps, hypothesis counts,
blend modes, cross-check on and off) gives, pair by pair, exactly what the
plain per-pair runner gives for that pair alone - keypoints, descriptors,
matches, homography, inlier mask, canvas and mask bit for bit (the plain path
itself is pinned to the oracle chain in test_gpu_pipeline). The result does
not depend on graph replay, on the launch shapes the autotuner picks, or on
what else is in the batch.

Served: a request's result is the same alone and inside a mixed arrival
trace, and matches the plain adaptive runner for that pair.
"""
import json

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SIZES = [(640, 480), (1280, 720), (1920, 1080), (800, 600)]


def make_pairs(n, seed0=4100, sizes=SIZES, overlap=(0.3, 0.6)):
    from cs_stitch import pipeline as P
    from cs_stitch import stages as st
    pairs = []
    for i in range(n):
        w, h = sizes[i % len(sizes)]
        a, b, H = P.synth_pair(seed0 + i, w, h, overlap=overlap)
        ta, tb = st.pitched(h, w, 3), st.pitched(h, w, 3)
        ta.copy_(torch.from_numpy(a).cuda())
        tb.copy_(torch.from_numpy(b).cuda())
        pairs.append((ta, tb))
    torch.cuda.synchronize()
    return pairs


def knobs(i):
    blends = ["multiband_6", "feather", "multiband_5", "multiband_3"]
    return dict(S=1.0 if i % 3 else 0.5, K=[500, 1500, 3000, 800][i % 4], R=[256, 1024, 512][i % 3],
                blend=blends[i % 4], cross_check=bool(i % 2), seed=11 * i)


def download_outcome(o, K):
    """Host copies of every output of a batch outcome (or runner outputs)."""
    from cs_stitch import pipeline as P
    na, nb = o["na"], o["nb"]
    d = {}
    for key, kp, n in (("kp_a", o["ka"], na), ("kp_b", o["kb"], nb)):
        d[key] = np.stack([P.download(kp.x, 4 * n).view(np.float32), P.download(kp.y, 4 * n).view(np.float32),
                           P.download(kp.response, 4 * n).view(np.float32)], 1)
    d["desc_a"] = P.download(o["da"].data, 32 * na)
    d["desc_b"] = P.download(o["db"].data, 32 * nb)
    nm = P.download(o["matches"].d_count, 4).view(np.int32)[0]
    d["nm"] = nm
    d["idx_a"] = P.download(o["matches"].idx_a, 4 * nm)
    d["idx_b"] = P.download(o["matches"].idx_b, 4 * nm)
    d["dist"] = P.download(o["matches"].distance, 4 * nm)
    d["H"] = np.array(o["h"].H[:])
    d["h_status"] = o["h"].status
    d["inliers"] = o["h"].num_inliers
    d["mask_inl"] = P.download(o["inlier_mask"], nm)
    if o.get("canvas") is not None and o["canvas"].data:
        d["canvas"] = P.download_image(o["canvas"], 3)
        d["mask"] = P.download_image(o["mask"], 1)
    return d


def outcome_dict(oc):
    return dict(na=oc.num_kp_a, nb=oc.num_kp_b, ka=oc.keypoints_a, kb=oc.keypoints_b, da=oc.descriptors_a,
                db=oc.descriptors_b, matches=oc.matches, h=oc.homography, inlier_mask=oc.inlier_mask,
                canvas=oc.canvas if oc.status == 0 else None, mask=oc.mask)


def run_plain(pairs, cfgs):
    from cs_stitch import pipeline as P
    W = max(t[0].shape[1] for t in pairs)
    H = max(t[0].shape[0] for t in pairs)
    runner = P.PipelineRunner(W, H, max_k=4000, max_r=2048, canvas_width=4096, canvas_height=4096,
                              workspace_mode=P.F.BLEND_NAMES["multiband_7"])
    res = []
    for (ta, tb), c in zip(pairs, cfgs):
        runner.use_device_pair(ta, tb)
        rec = runner.execute_run(P.harness_config(S=c["S"], K=c["K"], R=c["R"], blend=c["blend"],
                                                  cross_check=c["cross_check"], seed=c["seed"]))
        out = runner.outputs()
        o = dict(na=rec.num_kp_a, nb=rec.num_kp_b, ka=out["ka"], kb=out["kb"], da=out["da"], db=out["db"],
                 matches=out["matches"], h=out["h"], inlier_mask=out["inlier_mask"],
                 canvas=out["canvas"] if rec.status == b"OK" else None, mask=out["mask"])
        res.append((rec.status.decode(), download_outcome(o, c["K"])))
    runner.close()
    return res


def same(d1, d2, what):
    assert d1.keys() == d2.keys(), (what, d1.keys() ^ d2.keys())
    for k in d1:
        assert np.array_equal(np.asarray(d1[k]), np.asarray(d2[k])), (what, k)


def run_batch(br, pairs, cfgs):
    from cs_stitch import batch as B
    specs = [B.pair_spec(ta, tb, **c) for (ta, tb), c in zip(pairs, cfgs)]
    outs, t = br.run(specs)
    return [(o.status, download_outcome(outcome_dict(o), c["K"])) for o, c in zip(outs, cfgs)], t


def test_batch_matches_plain_runner_pair_by_pair():
    from cs_stitch import batch as B
    pairs = make_pairs(8)
    cfgs = [knobs(i) for i in range(8)]
    plain = run_plain(pairs, cfgs)
    br = B.BatchRunner(max_pairs=16, pool_bytes=8 << 30)
    got, t = run_batch(br, pairs, cfgs)
    print("batch ms", t.feature_ms, t.render_ms, t.host_ms, t.total_ms)
    for i, ((st_p, dp), (st_b, db)) in enumerate(zip(plain, got)):
        print(i, cfgs[i], st_p, st_b, dp["nm"], dp["inliers"], dp.get("canvas", np.zeros(0)).shape)
        assert (st_b == 0) == (st_p == "OK"), (i, st_p, st_b)
        same(dp, db, f"pair {i}")
    assert sum(s == "OK" for s, _ in plain) >= 6
    br.close()


def test_batch_independent_of_graphs_tuning_and_company():
    """Direct launches vs graph replay, other launch shapes, a different batch
    in between, and a pair alone vs inside a batch all give the same bits."""
    from cs_stitch import batch as B
    pairs = make_pairs(6, seed0=5200)
    cfgs = [knobs(i + 1) for i in range(6)]
    br = B.BatchRunner(max_pairs=8, pool_bytes=6 << 30)
    ref, _ = run_batch(br, pairs, cfgs)
    other = make_pairs(3, seed0=7300, sizes=[(1280, 720)])
    # (graph replay, S4 segment, workgroups per CU, per-stage workgroups S1..S6, S1 tile rows)
    shapes = ((False, 128, 8, (0,) * 6, 64), (True, 64, 4, (0,) * 6, 64), (True, 512, 16, (0,) * 6, 32),
              (True, 256, 8, (4, 32, 16, 8, 24, 4), 16), (False, 128, 8, (32, 4, 4, 32, 8, 16), 16),
              (True, 128, 8, (0,) * 6, 64))
    for graphs, seg, wpc, stages, rows in shapes:
        br.apply(s4_segment=seg, workgroups_per_cu=wpc, graphs=graphs, stage_workgroups=stages, s1_tile_rows=rows)
        run_batch(br, other, [knobs(0)] * 3)  # different work between replays
        got, _ = run_batch(br, pairs, cfgs)
        for i in range(6):
            assert got[i][0] == ref[i][0]
            same(ref[i][1], got[i][1], f"graphs={graphs} seg={seg} wpc={wpc} stages={stages} rows={rows} pair {i}")
    alone, _ = run_batch(br, pairs[2:3], cfgs[2:3])
    same(ref[2][1], alone[0][1], "alone")
    br.close()


def test_autotune_picks_and_caches(tmp_path):
    from cs_stitch import batch as B
    pairs = make_pairs(4, seed0=6100, sizes=[(640, 480), (1280, 720)])
    cfgs = [knobs(i) for i in range(4)]
    specs = [B.pair_spec(ta, tb, **c) for (ta, tb), c in zip(pairs, cfgs)]
    br = B.BatchRunner(max_pairs=4, pool_bytes=4 << 30)
    cache = tmp_path / "tune.json"
    t1 = br.autotune(specs, str(cache))
    assert t1.s4_segment in (64, 128, 256, 512) and t1.workgroups_per_cu in (8, 16, 24, 32)
    entries = json.loads(cache.read_text())
    assert len(entries) == 1
    (entry,) = entries.values()
    assert entry["s4_segment"] == t1.s4_segment
    br2 = B.BatchRunner(max_pairs=4, pool_bytes=4 << 30)
    t2 = br2.autotune(specs, str(cache))  # served from the cache
    assert (t2.s4_segment, t2.workgroups_per_cu) == (t1.s4_segment, t1.workgroups_per_cu)
    br.close()
    br2.close()


def test_batch_rejects_what_does_not_fit():
    """A canvas whose pyramid exceeds the pool fails alone; the rest run."""
    from cs_stitch import batch as B
    pairs = make_pairs(2, seed0=8100, sizes=[(640, 480)])
    cfgs = [dict(knobs(1), blend="multiband_7"), dict(knobs(1), blend="feather")]
    pool = 48 << 20
    br = B.BatchRunner(max_pairs=2, pool_bytes=pool)
    specs = [B.pair_spec(ta, tb, **c) for (ta, tb), c in zip(pairs, cfgs)]
    outs, _ = br.run(specs)
    # A slot holds half the pool: a quarter of it for pairs, the rest for canvases.
    slot = pool // 2
    assert 2 * B.pair_bytes(640, 480, cfgs[0]["K"], cfgs[0]["R"]) <= slot // 4
    g0, g1 = outs[0].geometry, outs[1].geometry
    assert B.render_bytes(g0.width, g0.height, "multiband_7") > slot - slot // 4
    assert B.render_bytes(g1.width, g1.height, "feather") <= slot - slot // 4
    assert outs[0].status == 2 and outs[0].fail_stage == b"s6"
    assert outs[1].status == 0 and outs[1].canvas.data
    cap, used, peak = br.pool()
    assert cap >= pool - 4096 and peak <= cap
    br.close()


def test_batch_stream_equals_batch_by_batch():
    """submit() overlaps a batch's S1-S5 with the previous batch's S6 on two
    streams and hands back the previous batch's outcomes: bit for bit what
    run() gives for each batch on its own."""
    from cs_stitch import batch as B
    groups = [make_pairs(4, seed0=9100 + 10 * g, sizes=SIZES[g % 2:] + SIZES[:g % 2]) for g in range(3)]
    cfgs = [[knobs(i + g) for i in range(4)] for g in range(3)]
    br = B.BatchRunner(max_pairs=4, pool_bytes=4 << 30)
    ref = [run_batch(br, groups[g], cfgs[g])[0] for g in range(3)]
    specs = [[B.pair_spec(ta, tb, **c) for (ta, tb), c in zip(groups[g], cfgs[g])] for g in range(3)]
    got = []
    prev, _ = br.submit(specs[0])
    assert prev is None
    for g in (1, 2):
        prev, t = br.submit(specs[g])
        assert len(prev) == 4 and t.render_ms > 0
        got.append([(o.status, download_outcome(outcome_dict(o), c["K"])) for o, c in zip(prev, cfgs[g - 1])])
    last = br.finish()
    got.append([(o.status, download_outcome(outcome_dict(o), c["K"])) for o, c in zip(last, cfgs[2])])
    for g in range(3):
        for i in range(4):
            assert got[g][i][0] == ref[g][i][0], (g, i)
            same(ref[g][i][1], got[g][i][1], f"stream batch {g} pair {i}")
    assert sum(s == 0 for grp in ref for s, _ in grp) >= 9
    # run() after a stream completes whatever is pending and still agrees.
    br.submit(specs[0])
    again, _ = run_batch(br, groups[1], cfgs[1])
    for i in range(4):
        same(ref[1][i][1], again[i][1], f"run after submit pair {i}")
    br.close()


def test_probe_times_every_s6_kernel_without_changing_results():
    """The per-kernel probe (direct launches, events around each kernel)
    leaves every output bit unchanged and accounts for each launch: the warp
    once per render, one pull and one push per pyramid level, the level-0 mix;
    compulsory bytes follow the canvases' pixel counts."""
    from cs_stitch import batch as B
    pairs = make_pairs(4, seed0=9500)
    cfgs = [dict(knobs(i), blend=b) for i, b in enumerate(["multiband_6", "feather", "multiband_3", "multiband_6"])]
    br = B.BatchRunner(max_pairs=4, pool_bytes=4 << 30)
    ref, _ = run_batch(br, pairs, cfgs)
    br.set_probe(True)
    got, _ = run_batch(br, pairs, cfgs)
    got2, _ = run_batch(br, pairs, cfgs)
    stats = {(e["kernel"], e["level"]): e for e in br.probe_read()}
    br.set_probe(False)
    for i in range(4):
        same(ref[i][1], got[i][1], f"probed pair {i}")
        same(ref[i][1], got2[i][1], f"probed twice pair {i}")
    assert stats[("s6_warp", 0)]["launches"] == 2
    assert stats[("s6_reduce_gauss", 1)]["launches"] == 2
    assert stats[("s6_mix", 0)]["launches"] == 2
    for e in stats.values():
        assert e["total_ms"] > 0 and e["algorithmic_bytes"] > 0, e
    # The feather canvas alone makes the warp read A and B (3 B per pixel each)
    # and write the panorama and mask (4 B per canvas pixel).
    outs, _ = br.run([B.pair_spec(ta, tb, **c) for (ta, tb), c in zip(pairs, cfgs)])
    g = outs[1].geometry
    assert stats[("s6_warp", 0)]["algorithmic_bytes"] / 2 > 3 * 2 * 1280 * 720 + 4 * g.width * g.height
    br.close()


def serve_requests(pairs, arrivals, seeds, keep):
    from cs_stitch import batch as B
    return [B.request(i, ta, tb, t, seed=s, keep_output=k)
            for i, ((ta, tb), t, s, k) in enumerate(zip(pairs, arrivals, seeds, keep))]


def completion_bits(c):
    from cs_stitch import pipeline as P
    d = dict(status=c.status, attempts=c.attempts, knobs=(c.S, c.K, c.R, c.blend_mode), H=np.array(c.homography.H[:]),
             inliers=c.homography.num_inliers, matches=c.homography.num_matches, kp=(c.num_kp_a, c.num_kp_b))
    if c.status == b"OK":
        d["canvas"] = P.download_image(c.canvas, 3)
        d["mask"] = P.download_image(c.mask, 1)
        d["stats"] = (c.statistics.union_pixels, c.statistics.overlap_pixels)
    return d


def test_server_result_same_alone_and_in_the_mix():
    from cs_stitch import batch as B
    from cs_stitch import pipeline as P
    sizes = [(640, 480), (1280, 720), (800, 600), (1920, 1080)]
    pairs = make_pairs(8, seed0=9100, sizes=sizes, overlap=(0.5, 0.8))
    srv = B.Server(hbm_budget=16 << 30, max_batch=8)
    target = 3
    alone, st = srv.run_trace(serve_requests([pairs[target]], [0.0], [5], [True]))
    a_bits = completion_bits(alone[0])
    srv.release(0)
    # Dense arrivals (a burst, then a few more while the first are on the GPU).
    arrivals = [0.0, 0.0, 0.0, 0.0] + [0.2 + 0.1 * k for k in range(4)]
    mix, st = srv.run_trace(serve_requests(pairs, arrivals, [5] * 8, [i == target for i in range(8)]))
    print("served", st.completed, st.failed, st.p50_ms, st.p99_ms, st.feature_launches, st.mean_feature_batch,
          st.render_launches, st.mean_render_batch, [c.status for c in mix])
    assert len(mix) == 8 and st.completed + st.failed == 8
    assert st.mean_feature_batch > 1.0  # requests really shared launches
    m = [c for c in mix if c.id == target][0]
    b_bits = completion_bits(m)
    same(a_bits, b_bits, "alone vs mix")
    # Same pair through the plain adaptive runner.
    ta, tb = pairs[target]
    runner = P.PipelineRunner(ta.shape[1], ta.shape[0], max_k=4000, max_r=2048, canvas_width=4096,
                              canvas_height=4096)
    runner.use_device_pair(ta, tb)
    rec = runner.execute_adaptive_run(P.Controller(), seed=5)
    assert rec.status == m.status and rec.attempt_count == m.attempts
    assert (rec.S, rec.K, rec.R) == (m.S, m.K, m.R)
    if rec.status == b"OK":
        assert np.array_equal(np.array(rec.H[:]), b_bits["H"])
        out = runner.outputs()
        assert np.array_equal(P.download_image(out["canvas"], 3), b_bits["canvas"])
    srv.release(target)
    runner.close()
    srv.close()


def test_server_fixed_knobs_matches_batch():
    """Non-adaptive serving renders every request; each canvas equals the
    batched runner's for the same pair and knobs, alone or mixed."""
    from cs_stitch import batch as B
    pairs = make_pairs(6, seed0=9500, sizes=[(640, 480), (1280, 720), (1920, 1080)])
    cfgs = [dict(knobs(i), seed=3) for i in range(6)]
    br = B.BatchRunner(max_pairs=6, pool_bytes=8 << 30)
    ref, _ = run_batch(br, pairs, cfgs)
    br.close()
    srv = B.Server(hbm_budget=12 << 30, max_batch=4, adaptive=False)
    reqs = [B.request(i, ta, tb, 0.5 * i, seed=3, fixed=B.pair_spec(ta, tb, **c), keep_output=True)
            for i, ((ta, tb), c) in enumerate(zip(pairs, cfgs))]
    done, st = srv.run_trace(reqs)
    assert st.completed == sum(r[0] == 0 for r in ref)
    for c in done:
        i = c.id
        if ref[i][0] != 0:
            continue
        from cs_stitch import pipeline as P
        assert np.array_equal(P.download_image(c.canvas, 3), ref[i][1]["canvas"]), i
        assert np.array_equal(np.array(c.homography.H[:]), ref[i][1]["H"]), i
        srv.release(i)
    srv.close()


@pytest.mark.parametrize("kind", ["v3", "v4"])
def test_server_learned_policy_matches_plain_runner(kind):
    """Served requests under a learned policy (scene features after the
    batched probe, gate features after each batched S5) end exactly as the
    plain adaptive runner under the same policy."""
    import json
    from cs_stitch import batch as B
    from cs_stitch import pipeline as P
    from test_policies import gated_v3_policy, gated_v4_policy
    text = json.dumps(gated_v3_policy() if kind == "v3" else gated_v4_policy())
    pairs = make_pairs(4, seed0=9300, sizes=[(640, 480), (800, 600)], overlap=(0.5, 0.8))
    srv = B.Server(hbm_budget=12 << 30, max_batch=4, policy_json=text)
    done, st = srv.run_trace(serve_requests(pairs, [0.0] * 4, [5] * 4, [False] * 4))
    assert len(done) == 4 and st.completed + st.failed == 4
    runner = P.PipelineRunner(800, 600, max_k=4000, max_r=2048, canvas_width=4096, canvas_height=4096)
    for c in done:
        ta, tb = pairs[c.id]
        runner.use_device_pair(ta, tb)
        rec = runner.execute_adaptive_run(P.Controller(policy_json=text), seed=5)
        print(kind, c.id, c.status, c.attempts, (c.S, c.K, c.R))
        assert rec.status == c.status and rec.attempt_count == c.attempts
        assert (rec.S, rec.K, rec.R) == (c.S, c.K, c.R)
        if rec.status == b"OK":
            assert np.array_equal(np.array(rec.H[:]), np.array(c.homography.H[:]))
    runner.close()
    srv.close()
    with pytest.raises(Exception):
        B.Server(hbm_budget=1 << 30, policy_json='{"schema_version": "controller-policy-v9"}')
