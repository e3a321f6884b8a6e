"""Controller observations: scene features (before matching) and gate
features (after S5), src/controller/features.cu:11-171 and
src/controller/relational_features.cu:6-190.

CPU tests pin the oracle's restatement with closed forms that do not reuse
its formulas: constant, ramp and checkerboard statistics, an exact
translation for the shift search, identical images under H = I for the
photometric and patch terms, hand-counted inlier grids, residuals and match
distances. GPU tests compare the HIP kernels with the oracle on the same
inputs: relative difference at most 1e-12 (absolute 1e-12 near zero) for
every feature. Bit agreement with the reference's CUDA build is not claimed:
nvcc contracts multiply-adds by default, and log and hypot come from
different math libraries.
"""
import math

import numpy as np
import pytest

import images
import oracle_lib as O

RTOL, ATOL = 1e-12, 1e-12


def tex_rgb(w, h, seed):
    return images.texture_rgb(seed, h, w)


def tex_gray(w, h, seed):
    return np.ascontiguousarray(images.gray_of(tex_rgb(w, h, seed)), np.uint8)


def ramp(w, h):
    return np.tile(np.arange(w, dtype=np.uint8), (h, 1))


def checker(w, h):
    yy, xx = np.mgrid[0:h, 0:w]
    return ((xx + yy) % 2 * 255).astype(np.uint8)


# ---------------------------------------------------------------- oracle, scene

def test_scene_constant_ramp_and_checkerboard():
    w, h = 64, 48
    st, f = O.scene_features(np.full((h, w), 51, np.uint8), ramp(w, h), relations=False)
    assert st == 0
    assert f[0] == 0.0  # constant A: no gradient
    for q in range(4):
        assert f[2 + 3 * q] == pytest.approx(0.2, rel=1e-14)
        assert f[3 + 3 * q] < 1e-7 and f[4 + 3 * q] == 0.0
    # B = x / 255: horizontal steps of 1/255 on every row, none vertically.
    assert f[1] == pytest.approx(h * (w - 1) / 255 / (h * (w - 1) + w * (h - 1)), rel=1e-13)
    half = np.arange(32)
    for q, xs in enumerate((half, half + 32, half, half + 32)):
        assert f[14 + 3 * q] == pytest.approx(xs.mean() / 255, rel=1e-13)
        assert f[15 + 3 * q] == pytest.approx(xs.std() / 255, rel=1e-9)
        assert f[16 + 3 * q] == pytest.approx(24 * 31 / 255 / (24 * 31 + 32 * 23), rel=1e-13)
    assert np.all(f[26:] == 0)  # relations off
    st, f = O.scene_features(checker(w, h), checker(w + 4, h + 4), relations=False)
    assert st == 0 and f[0] == 1.0 and f[1] == 1.0
    for q in range(8):
        assert f[2 + 3 * q] == pytest.approx(0.5, abs=1e-15)
        assert f[3 + 3 * q] == pytest.approx(0.5, abs=1e-12)
        assert f[4 + 3 * q] == 1.0


def test_scene_statistics_agree_with_plain_sums():
    rng = np.random.default_rng(11)
    a = rng.integers(0, 256, (37, 53), dtype=np.uint8)
    b = tex_gray(61, 45, 9)
    st, f = O.scene_features(a, b, relations=False)
    assert st == 0
    for img, base in ((a, 2), (b, 14)):
        h, w = img.shape
        g = img / 255.0
        for q in range(4):
            x0, y0 = (q % 2) * w // 2, (q // 2) * h // 2
            x1, y1 = (q % 2 + 1) * w // 2, (q // 2 + 1) * h // 2
            t = g[y0:y1, x0:x1]
            grad = np.abs(np.diff(t, axis=1)).sum() + np.abs(np.diff(t, axis=0)).sum()
            edges = t.shape[0] * (t.shape[1] - 1) + t.shape[1] * (t.shape[0] - 1)
            assert f[base + 3 * q] == pytest.approx(t.mean(), rel=1e-13)
            assert f[base + 3 * q + 1] == pytest.approx(t.std(), rel=1e-10)
            assert f[base + 3 * q + 2] == pytest.approx(grad / edges, rel=1e-13)


def test_pair_relations_identity_and_translation():
    # Identical textured images: the peak is the zero shift with NCC 1 and
    # full overlap.
    g = tex_gray(97, 73, 4)
    st, f = O.scene_features(g, g)
    assert st == 0
    assert f[26] == pytest.approx(1.0, abs=1e-12) and f[30] == f[26]
    assert (f[28], f[29], f[31]) == (0.0, 0.0, 1.0)
    assert 0 < f[27] < 1
    # B(X + 8) = A(X) on a width-65 pair: the shift dx = +1/8 of (w - 1) is
    # exact, its NCC is 1 and 28 of the 32 grid columns stay inside B.
    wide = tex_gray(73, 49, 5)
    a, b = wide[:, 8:], wide[:, :65]
    st, f = O.scene_features(np.ascontiguousarray(a), np.ascontiguousarray(b))
    assert st == 0
    assert (f[28], f[29]) == (0.125, 0.0)
    assert f[26] > 0.999999 and f[31] == 28 * 32 / 1024


def test_scene_rejects_rgb_and_tiny_images():
    g = np.zeros((8, 8), np.uint8)
    assert O.scene_features(np.zeros((8, 8, 3), np.uint8), g)[0] == 1
    assert O.scene_features(np.zeros((8, 1), np.uint8), g)[0] == 1


# ---------------------------------------------------------------- oracle, gate

def identity_scale(w, h):
    st, si = O.scale_info(w, h, 1.0)
    assert st == 0
    return si


def gate_case(w=96, h=64):
    """Identical RGB images, H = I, scale 1. Ten matches: nine inliers (four
    in grid bin 0, three in bin 5, two in bin 15) and one outlier."""
    img = tex_rgb(w, h, 3)
    pts = [(3, 2), (10, 5), (20, 12), (5, 14), (30, 20), (40, 25), (35, 30), (80, 55), (90, 60), (60, 40)]
    kp_a = np.array(pts, np.float32)
    kp_b = kp_a.copy()
    idx = np.arange(10, dtype=np.int32)
    mask = np.array([1] * 9 + [0], np.uint8)
    dist = np.array([5, 5, 5, 5, 5, 5, 5, 5, 100, 200], np.float32)
    forward = [(10, 20, 0, 1), (18, 20, 1, 2), (0, 0, 2, 3), (5, math.inf, 1, -1)] + [(math.inf, math.inf, -1, -1)] * 6
    return img, kp_a, kp_b, idx, mask, dist, forward


def test_gate_identical_images_closed_forms():
    w, h = 96, 64
    img, kp_a, kp_b, idx, mask, dist, forward = gate_case(w, h)
    si = identity_scale(w, h)
    st, f = O.gate_features(img, img, kp_a, kp_b, idx, idx, dist, mask, 10, 9, 0.9, 0.4, 0.25, np.eye(3), si, si,
                            forward)
    assert st == 0
    assert list(f[:6]) == [10 / 4000, 9 / 4000, 0.9, 0.4, 0.25, 0.25]
    assert f[6] == pytest.approx((w - 1) * (h - 1) / (w * h), rel=1e-15)
    assert (f[7], f[8], f[9]) == ((w - 1) / w, (h - 1) / h, 1.0)
    # Photometric: everything overlaps and agrees.
    assert f[10] == 1.0 and f[11] == 0.0 and f[12] == 0.0 and f[14] == 0.0 and f[17] == 0.0
    assert f[13] == pytest.approx(1.0, abs=1e-12) and f[15] == f[16] > 0.05
    # Inlier spread: bins (4, 3, 2) of 9 in both images.
    p = np.array([4, 3, 2]) / 9
    for k in (0, 1):
        assert f[18 + k] == 3 / 16 and f[20 + k] == 4 / 9
        assert f[22 + k] == pytest.approx(-(p * np.log(p)).sum() / np.log(16), rel=1e-14)
    # Relations: zero residuals, distances 8 x 5, 100, 200 (the 90th
    # percentile is the 9th sorted value), ratios 0.5, 0.9, 1 (second 0).
    assert list(f[24:28]) == [0.0, 0.0, 0.0, 0.0]
    assert f[28] == pytest.approx((8 * 5 + 100 + 200) / 10 / 256, rel=1e-15)
    assert f[29] == 100 / 256
    assert f[30] == pytest.approx(0.8, rel=1e-15) and f[31] == pytest.approx(2 / 3, rel=1e-15)
    # Patches: identical, every patch fully usable.
    assert f[32] == pytest.approx(1.0, abs=1e-12) and f[33] == pytest.approx(1.0, abs=1e-12)
    assert f[34] == 0.0 and f[35] == pytest.approx(1.0, abs=1e-12) and f[36] == 1.0 and f[37] == 0.0


def test_gate_residual_percentiles_and_regions():
    w, h = 96, 64
    # Inliers 0, 1, 4, 5 sit in tiles 0, 0, 5, 5 (x in quarters of 96, y of
    # 64); B's copies are shifted by 0.5, 1, 2 and 40 px.
    img, kp_a, kp_b, idx, mask, dist, forward = gate_case(w, h)
    rows = [0, 1, 4, 5]
    kp_b[rows, 0] += np.array([0.5, 1.0, 2.0, 40.0], np.float32)
    mask = np.zeros(10, np.uint8)
    mask[rows] = 1
    si = identity_scale(w, h)
    st, f = O.gate_features(img, img, kp_a, kp_b, idx, idx, dist, mask, 10, 4, 0.4, 0.2, 1.0, np.eye(3), si, si,
                            forward, relation_mask=2)
    assert st == 0
    # Residual bins 4, 8, 16 and 256 (clamped at 32 px): median bin 8, 90th 256.
    assert (f[24], f[25]) == (1.0, 32.0)
    tiles = {}
    for i, e in zip(rows, [0.5, 1.0, 2.0, 40.0]):
        tx, ty = int(kp_a[i, 0] * 4 / w), int(kp_a[i, 1] * 4 / h)
        tiles.setdefault(tx + 4 * ty, []).append(e)
    means = np.array([np.mean(v) for v in tiles.values()])
    assert sorted(tiles) == [0, 5]
    assert f[26] == means.max() and f[27] == pytest.approx(means.std(), rel=1e-12)
    # Matching group off: no distances or ratios are seen, so the distance
    # mean is 0 and the 90th percentile, ratio mean and ambiguous share take
    # their empty values (1); the photometric group stays zero.
    assert list(f[28:32]) == [0.0, 1.0, 1.0, 1.0] and list(f[32:38]) == [0.0] * 6


def test_gate_invalid_inputs():
    w, h = 96, 64
    img, kp_a, kp_b, idx, mask, dist, forward = gate_case(w, h)
    si = identity_scale(w, h)
    args = (img, img, kp_a, kp_b, idx, idx, dist, mask)
    assert O.gate_features(*args, 10, 9, .9, .4, .25, np.zeros((3, 3)), si, si, forward)[0] == 1  # singular H
    assert O.gate_features(*args, 3, 3, .9, .4, .25, np.eye(3), si, si, forward)[0] == 1  # fewer than 4
    assert O.gate_features(*args, 11, 9, .9, .4, .25, np.eye(3), si, si, forward)[0] == 1  # above capacity
    big = np.zeros((4001, 2), np.float32)
    big[:10] = kp_a
    st, _ = O.gate_features(img, img, big, kp_b, idx, idx, dist, mask, 10, 9, .9, .4, .25, np.eye(3), si, si, None)
    assert st == 3  # the relations' keypoint limit
    st, f = O.gate_features(img, img, big, kp_b, idx, idx, dist, mask, 10, 9, .9, .4, .25, np.eye(3), si, si, None,
                            relation_mask=0)
    assert st == 0 and np.all(f[24:] == 0)


# ---------------------------------------------------------------- GPU parity

# Gate values computed through log (the inlier entropies, 22-23) or hypot
# (the alignment-relation residual statistics, 24-31): the device math
# library may round these differently from the host's, so they are held to
# the tolerance. Every other value is a fixed sequence of correctly rounded
# operations (sums in the reference's partial order, sqrt, divisions) and must
# be bit-equal.
GATE_LIBM = [22, 23] + list(range(24, 32))


def assert_features_close(got, want):
    got, want = np.asarray(got), np.asarray(want)
    exact = np.ones(got.shape, bool)
    if got.shape == (38,):
        exact[GATE_LIBM] = False
    np.testing.assert_array_equal(got[exact], want[exact])
    np.testing.assert_allclose(got[~exact], want[~exact], rtol=RTOL, atol=ATOL)


@pytest.mark.gpu
@pytest.mark.parametrize("shape_a,shape_b,kind", [((120, 160), (120, 160), "texture"), ((67, 93), (71, 88), "noise"),
                                                  ((2, 2), (3, 5), "noise"), ((270, 480), (270, 480), "texture"),
                                                  ((64, 80), (64, 80), "constant")])
@pytest.mark.parametrize("relations", [True, False])
def test_extract_scene_matches_oracle(shape_a, shape_b, kind, relations):
    import torch
    from cs_stitch import stages
    rng = np.random.default_rng(shape_a[0] * 31 + shape_b[1])
    if kind == "texture":
        a = tex_gray(shape_a[1], shape_a[0], 1)
        b = tex_gray(shape_b[1], shape_b[0], 2)
    elif kind == "constant":
        # Zero variances: the NCC terms are rounding residues over the 1e-12
        # floor, so only identical arithmetic agrees here.
        a = np.full(shape_a, 77, np.uint8)
        b = np.full(shape_b, 200, np.uint8)
    else:
        a = rng.integers(0, 256, shape_a, dtype=np.uint8)
        b = rng.integers(0, 256, shape_b, dtype=np.uint8)
    ga = stages.pitched(*shape_a, 1, pad=64)
    gb = stages.pitched(*shape_b, 1, pad=13)
    ga.copy_(torch.from_numpy(a).cuda())
    gb.copy_(torch.from_numpy(b).cuda())
    st, got = stages.extract_scene(ga, gb, relations)
    ost, want = O.scene_features(a, b, relations)
    assert st == ost == 0
    print(shape_a, shape_b, kind, relations, "bit-equal:", int((got == want).sum()), "of 32")
    assert_features_close(got, want)


def device_gate_chain(torch, st, a, b, S, K, R, cross):
    """S1-S5 through the stage wrappers; returns everything extract_gate reads."""
    ra, rb = (st.pitched(*x.shape[:2], 3) for x in (a, b))
    ra.copy_(torch.from_numpy(a).cuda())
    rb.copy_(torch.from_numpy(b).cuda())
    ga, gb, sa, sb = st.s1_preprocess(ra, rb, S)
    ka, kb = st.s2_detect(ga, 20, K), st.s2_detect(gb, 20, K)
    da, db = st.s3_describe(ga, ka, capacity=K), st.s3_describe(gb, kb, capacity=K)
    m, ws = st.s4_match(da, ka.count, db, kb.count, 0.8, cross, capacity_a=K, capacity_b=K)
    s5 = st.S5Runner(R, K)
    h = s5.estimate(ka, kb, m, sa, sb, R, capacity_a=K, capacity_b=K)
    return dict(ra=ra, rb=rb, ka=ka, kb=kb, da=da, db=db, m=m, ws=ws, s5=s5, h=h, sa=sa, sb=sb)


def oracle_scale(si):
    out = O.ScaleInfo()
    for name, _ in O.ScaleInfo._fields_:
        setattr(out, name, getattr(si, name))
    return out


def host_forward(d, K):
    """The device forward top-2 view, checked against the oracle's top-2."""
    from cs_stitch import stages
    raw = d["ws"][: K * 24].cpu().numpy().tobytes()
    arr = (stages.F.Top2 * K).from_buffer_copy(raw)
    fwd = [(r.first_distance, r.second_distance, r.first_index, r.second_index) for r in arr]
    na, nb = int(d["ka"].count.item()), int(d["kb"].count.item())
    ref = O.s4_top2(d["da"][:na].cpu().numpy(), d["db"][:nb].cpu().numpy())
    assert fwd[:na] == ref and all(t[2] == -1 and t[3] == -1 for t in fwd[na:])
    return fwd


@pytest.mark.gpu
@pytest.mark.parametrize("seed,w,h,S,K,R,cross", [(41, 640, 480, 0.5, 1000, 256, False),
                                                  (42, 1280, 720, 0.75, 3000, 512, True)])
def test_extract_gate_matches_oracle(seed, w, h, S, K, R, cross):
    import torch
    from cs_stitch import pipeline as P
    from cs_stitch import stages as st
    a, b, _ = P.synth_pair(seed, w, h, overlap=(0.5, 0.7))
    d = device_gate_chain(torch, st, a, b, S, K, R, cross)
    hres = d["h"]
    assert hres.status == 0 and hres.num_matches >= 4
    nm = hres.num_matches
    fwd = host_forward(d, K)
    kpa = np.stack([d["ka"].x.cpu().numpy(), d["ka"].y.cpu().numpy()], 1)[:K]
    kpb = np.stack([d["kb"].x.cpu().numpy(), d["kb"].y.cpu().numpy()], 1)[:K]
    ia, ib, dist = (t[:nm].cpu().numpy() for t in (d["m"].idx_a, d["m"].idx_b, d["m"].distance))
    mask = d["s5"].mask[:nm].cpu().numpy()
    fptr = d["ws"].data_ptr()  # the forward top-2 view starts the S4 workspace
    for relation_mask in (0, 1, 2, 4, 7):
        status, got = st.extract_gate(d["ra"], d["rb"], d["ka"].c(K), d["kb"].c(K), d["m"].c(), d["s5"].mask.data_ptr(),
                                      hres, d["sa"], d["sb"], fptr, relation_mask)
        ost, want = O.gate_features(a, b, kpa, kpb, ia, ib, dist, mask, nm, hres.num_inliers, hres.inlier_ratio,
                                    hres.inlier_coverage, hres.reproj_residual, np.array(hres.H[:]),
                                    oracle_scale(d["sa"]), oracle_scale(d["sb"]), fwd, relation_mask,
                                    capacity=d["m"].idx_a.numel())
        assert status == ost == 0, (status, ost)
        print(seed, relation_mask, "bit-equal:", int((got == want).sum()), "of 38")
        assert_features_close(got, want)
    # Fewer than 4 matches: INVALID_ARGUMENT, as the oracle says.
    few = P.F.HomographyResult.from_buffer_copy(bytes(hres))
    few.num_matches = 3
    assert st.extract_gate(d["ra"], d["rb"], d["ka"].c(K), d["kb"].c(K), d["m"].c(), d["s5"].mask.data_ptr(), few,
                           d["sa"], d["sb"], fptr, 7)[0] == P.F.INVALID_ARGUMENT
    # A singular homography is refused before anything is enqueued.
    bad = P.F.HomographyResult.from_buffer_copy(bytes(hres))
    for i in range(9):
        bad.H[i] = 0.0
    assert st.extract_gate(d["ra"], d["rb"], d["ka"].c(K), d["kb"].c(K), d["m"].c(), d["s5"].mask.data_ptr(), bad,
                           d["sa"], d["sb"], fptr, 7)[0] == P.F.INVALID_ARGUMENT


@pytest.mark.gpu
def test_runner_collects_features():
    """The plain runner's observations: scene features of the pair at
    S = 0.25, gate features of the run, both equal to the oracle on the same
    data; collection off leaves them absent."""
    from cs_stitch import pipeline as P
    w, h, S, K = 640, 480, 0.5, 1000
    a, b, _ = P.synth_pair(77, w, h, overlap=(0.5, 0.7))
    runner = P.PipelineRunner(w, h, max_k=4000, max_r=2048, canvas_width=2048, canvas_height=1024)
    runner.upload_pair(a, b)
    runner.execute_run(P.harness_config(S=S, K=K, R=256))
    scene, gate, present, _ = runner.features()
    assert not present and not scene.any() and not gate.any()
    runner.set_collect_features(True)
    runner.upload_pair(a, b)
    rec = runner.execute_run(P.harness_config(S=S, K=K, R=256))
    assert rec.status == b"OK"
    scene, gate, present, t = runner.features()
    assert present and t > 0
    st0, want_scene = O.scene_features(O.s1_preprocess(a, 0.25), O.s1_preprocess(b, 0.25))
    assert st0 == 0
    assert_features_close(scene, want_scene)
    assert_features_close(gate, oracle_gate_of_runner(runner, rec, a, b, S, K, 7))
    runner.close()


def oracle_gate_of_runner(runner, rec, a, b, S, K, relation_mask, max_k=4000):
    """The oracle's gate features of the runner's last S5, from its device
    outputs (keypoints, matches, mask, homography); the forward top-2 view is
    recomputed on the host from its descriptors."""
    from cs_stitch import pipeline as P
    out = runner.outputs()
    hres = out["h"]
    nm = hres.num_matches
    kp = {}
    for key in ("ka", "kb"):
        kp[key] = np.stack([P.download(out[key].x, 4 * K).view(np.float32),
                            P.download(out[key].y, 4 * K).view(np.float32)], 1)
    m = out["matches"]
    ia = P.download(m.idx_a, 4 * nm).view(np.int32)
    ib = P.download(m.idx_b, 4 * nm).view(np.int32)
    dist = P.download(m.distance, 4 * nm).view(np.float32)
    mask = P.download(out["inlier_mask"], nm)
    fwd = None
    if relation_mask & 1:
        na, nb = (int(P.download(out[k].d_count, 4).view(np.int32)[0]) for k in ("ka", "kb"))
        da = P.download(out["da"].data, 32 * na).reshape(-1, 32)
        db = P.download(out["db"].data, 32 * nb).reshape(-1, 32)
        fwd = O.s4_top2(da, db) + [(math.inf, math.inf, -1, -1)] * (K - na)
    _, sa = O.scale_info(a.shape[1], a.shape[0], S)
    _, sb = O.scale_info(b.shape[1], b.shape[0], S)
    st, want = O.gate_features(a, b, kp["ka"], kp["kb"], ia, ib, dist, mask, nm, hres.num_inliers, hres.inlier_ratio,
                               hres.inlier_coverage, hres.reproj_residual, np.array(hres.H[:]), sa, sb, fwd,
                               relation_mask, capacity=max_k)
    assert st == 0
    return want


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["v3", "v4"])
def test_adaptive_runner_with_learned_policy(kind):
    """Adaptive runs under learned policies: the probe collects the scene
    features (pair relations when the pooled policy asks), the first rung is
    the controller's choice from them, gate features of the last attempt match
    the oracle on the runner's own outputs, and the run ends in a terminal
    status of the decision logic."""
    import json
    from cs_stitch import pipeline as P
    from test_policies import gated_v3_policy, gated_v4_policy
    ctl = P.Controller(policy_json=json.dumps(gated_v3_policy() if kind == "v3" else gated_v4_policy()))
    learned, pair_relations, relation_mask, _ = ctl.observations()
    assert learned and pair_relations == (kind == "v4")
    w, h = 640, 480
    a, b, _ = P.synth_pair(91, w, h, overlap=(0.5, 0.7))
    runner = P.PipelineRunner(w, h, max_k=4000, max_r=2048, canvas_width=2048, canvas_height=1024)
    runner.upload_pair(a, b)
    rec = runner.execute_adaptive_run(ctl, seed=0)
    assert rec.status.decode() in ("OK", "BUDGET_EXHAUSTED", "REJECTED"), (rec.status, rec.fail_stage)
    scene, gate, present, _ = runner.features()
    st, want = O.scene_features(O.s1_preprocess(a, 0.25), O.s1_preprocess(b, 0.25), pair_relations)
    assert st == 0
    assert_features_close(scene, want)
    assert (rec.probe_gradient_a, rec.probe_gradient_b) == (scene[0], scene[1])
    trace = runner.attempt_trace()
    print(kind, rec.status, [(t.S, t.K, t.R, t.action, t.reason) for t in trace], present)
    first = ctl.walk(scene[0], scene[1], w, h, w, h, [], [], scene=scene)[0]
    assert (trace[0].S, trace[0].K, trace[0].R) == first[2:]
    if present:
        last = trace[-1]
        assert_features_close(gate, oracle_gate_of_runner(runner, rec, a, b, last.S, last.K, relation_mask))
    runner.close()
