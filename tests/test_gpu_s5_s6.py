"""GPU parity of S5 (MSAC) and S6 (warp + blend) against the CPU oracle.

S5 bar: identical status, inlier mask and telemetry, H and H_s bit-exact
(the FP64 arithmetic and every summation order are the reference's), and the
known-answer recovery of tests/test_s5_cuda.cu:99-116 within 1e-3.
S6 bar: mask and render statistics exact, RGB within 1 LSB of the oracle
(the level-1 Gaussian is summed in another order) with differing pixels
rare, and golden vectors of the reference's FP64 renderer within 1 LSB.
"""
import ctypes as C

import numpy as np
import pytest

import images
import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    import torch
    from cs_stitch import stages
    stages.F.load()
    return torch, stages


def kpset(torch, st, xy, cap=None):
    n = xy.shape[0]
    cap = n if cap is None else cap
    k = st.KeypointSet.empty(cap)
    if n:
        k.x[:n] = torch.from_numpy(xy[:, 0].astype(np.float32)).cuda()
        k.y[:n] = torch.from_numpy(xy[:, 1].astype(np.float32)).cuda()
    k.count.fill_(n)
    return k


def matchset(torch, st, ia, ib, cap=None):
    n = len(ia)
    cap = n if cap is None else cap
    m = st.MatchSet.empty(cap)
    if n:
        m.idx_a[:n] = torch.from_numpy(np.asarray(ia, np.int32)).cuda()
        m.idx_b[:n] = torch.from_numpy(np.asarray(ib, np.int32)).cuda()
    m.count.fill_(n)
    return m


def fixture_points(n=80, outliers=True, line=False):
    """tests/test_s5_cuda.cu:69-85: projective warp, every 4th match an outlier."""
    x = 20 + (np.arange(n) % 10) * 50.0
    y = 2 * x if line else 20 + (np.arange(n) // 10) * 45.0
    x = x.astype(np.float32).astype(np.float64)
    y = y.astype(np.float32).astype(np.float64)
    z = 0.0002 * x - 0.0001 * y + 1
    u = ((1.02 * x + 0.03 * y + 11) / z).astype(np.float32)
    v = ((-0.02 * x + 0.98 * y + 7) / z).astype(np.float32)
    if outliers:
        i = np.arange(n)
        u = np.where(i % 4 == 0, (600 - i).astype(np.float32), u)
        v = np.where(i % 4 == 0, (400 - i * 2).astype(np.float32), v)
    return np.stack([u, v], 1), np.stack([x, y], 1)


def result_tuple(r):
    return (r.status, r.num_matches, r.num_inliers, r.inlier_ratio, r.inlier_coverage,
            np.float64(r.reproj_residual).tobytes(), np.array(r.H[:]).tobytes(), np.array(r.H_s[:]).tobytes())


def run_both(T, a_xy, b_xy, ia, ib, sa_wh, sb_wh, S, R, refine, thr, seed, runner=None, prev=0, gen=1, ostate=None):
    torch, st = T
    sa, sb = st.scale_info(*sa_wh, S), st.scale_info(*sb_wh, S)
    ka, kb = kpset(torch, st, a_xy), kpset(torch, st, b_xy)
    m = matchset(torch, st, ia, ib)
    runner = runner or st.S5Runner(max(R, 1), max(len(ia), 1))
    g = runner.estimate(ka, kb, m, sa, sb, R, previous=prev, batch_size=32, refinement_steps=refine, threshold=thr,
                        seed=seed, generation=gen)
    gmask = runner.mask[: len(ia)].cpu().numpy()
    _, osa = O.scale_info(*sa_wh, S)
    _, osb = O.scale_info(*sb_wh, S)
    o, omask, ostate = O.s5_estimate_kp(a_xy[:, 0], a_xy[:, 1], b_xy[:, 0], b_xy[:, 1], ia, ib, osa, osb, R, refine,
                                        thr, seed, previous_hypotheses=prev, state=ostate)
    return g, gmask, o, omask, runner, ostate


def test_s5_known_answer_recovery(T):
    a, b = fixture_points()
    idx = np.arange(80)
    g, gm, o, om, _, _ = run_both(T, a, b, idx, idx, (640, 480), (640, 480), 1.0, 257, 2, 0.5, 20261022)
    truth = np.array([1.02, 0.03, 11, -0.02, 0.98, 7, 0.0002, -0.0001, 1])
    assert g.status == 0 and g.num_matches == 80 and g.num_inliers == 60
    assert abs(g.inlier_ratio - 0.75) < 1e-12 and g.reproj_residual < 1e-3
    H = np.array(g.H[:])
    assert np.max(np.abs(H / H[8] - truth / truth[8])) < 1e-3
    assert np.array_equal(gm, (np.arange(80) % 4 != 0).astype(np.uint8))
    assert result_tuple(g) == result_tuple(o) and np.array_equal(gm, om)


def test_s5_split_r_equals_full_r(T):
    torch, st = T
    a, b = fixture_points()
    idx = np.arange(80)
    full, fmask, *_ = run_both(T, a, b, idx, idx, (640, 480), (640, 480), 1.0, 257, 2, 0.5, 20261022)
    runner = st.S5Runner(300, 80)
    g1, _, _, _, runner, ostate = run_both(T, a, b, idx, idx, (640, 480), (640, 480), 1.0, 17, 2, 0.5, 20261022,
                                           runner=runner)
    assert g1.status == 0
    # Same device buffers are required for an extension: reuse the runner's keypoints via a fresh call on the
    # same tensors is not possible through run_both, so extend through the stage API directly.
    sa, sb = st.scale_info(640, 480, 1.0), st.scale_info(640, 480, 1.0)
    ka, kb = kpset(torch, st, a), kpset(torch, st, b)
    m = matchset(torch, st, idx, idx)
    r17 = runner.estimate(ka, kb, m, sa, sb, 17, batch_size=32, refinement_steps=2, threshold=0.5, seed=20261022)
    ext = runner.estimate(ka, kb, m, sa, sb, 257, previous=17, batch_size=32, refinement_steps=2, threshold=0.5,
                          seed=20261022)
    assert r17.status == 0
    assert result_tuple(ext) == result_tuple(full)
    assert np.array_equal(runner.mask[:80].cpu().numpy(), fmask)
    # A changed generation makes the extension stale: INVALID_ARGUMENT in the result.
    stale = runner.estimate(ka, kb, m, sa, sb, 300, previous=257, batch_size=32, refinement_steps=2, threshold=0.5,
                            seed=20261022, generation=2)
    assert stale.status == 1


def test_s5_not_enough_points_and_bad_indices(T):
    a, b = fixture_points(3, outliers=False)
    g, _, o, _, _, _ = run_both(T, a, b, np.arange(3), np.arange(3), (640, 480), (640, 480), 1.0, 64, 1, 3.0, 0)
    assert g.status == 4 == o.status and g.num_matches == 3
    a, b = fixture_points(20, outliers=False)
    ia = np.arange(20)
    ib = ia.copy()
    ib[5] = 25  # out of range
    g, _, o, _, _, _ = run_both(T, a, b, ia, ib, (640, 480), (640, 480), 1.0, 64, 1, 3.0, 0)
    assert g.status == 1 == o.status


def test_s5_collinear_is_degenerate(T):
    a, b = fixture_points(40, outliers=False, line=True)
    idx = np.arange(40)
    g, gm, o, om, _, _ = run_both(T, a, b, idx, idx, (640, 480), (640, 480), 1.0, 128, 1, 3.0, 5)
    assert g.status == 5 == o.status and not gm.any()


@pytest.mark.parametrize("seed,m,R,S", [(1, 300, 512, 1.0), (2, 2000, 4096, 1.0), (3, 1200, 1024, 0.5),
                                        (4, 5000, 2048, 0.75)])
def test_s5_random_matches_against_oracle(T, seed, m, R, S):
    rng = np.random.default_rng(seed)
    W, Hh = 1920, 1080
    sw, sh = max(1, round(W * S)), max(1, round(Hh * S))
    bxy = np.stack([rng.uniform(0, sw - 1, m), rng.uniform(0, sh - 1, m)], 1).astype(np.float32)
    th = rng.uniform(-0.08, 0.08)
    Htrue = np.array([[np.cos(th), -np.sin(th), rng.uniform(-300, 300)],
                      [np.sin(th), np.cos(th), rng.uniform(-100, 100)], [rng.uniform(-1e-5, 1e-5), 0, 1]])
    p = np.c_[bxy, np.ones(m)] @ Htrue.T
    axy = (p[:, :2] / p[:, 2:]).astype(np.float32) + rng.normal(0, 0.7, (m, 2)).astype(np.float32)
    out = rng.random(m) < 0.35
    axy[out] = rng.uniform(0, sw, (out.sum(), 2)).astype(np.float32)
    # Keypoint arrays are larger than the match set; matches index into them.
    na, nb = m + 17, m + 9
    A = np.vstack([axy, rng.uniform(0, 100, (17, 2)).astype(np.float32)])
    B = np.vstack([bxy, rng.uniform(0, 100, (9, 2)).astype(np.float32)])
    ia, ib = np.arange(m), np.arange(m)
    g, gm, o, om, _, _ = run_both(T, A, B, ia, ib, (W, Hh), (W, Hh), S, R, 1, 3.0, seed)
    assert g.status == 0
    assert result_tuple(g) == result_tuple(o)
    assert np.array_equal(gm, om)


# ---------------------------------------------------------------- S6

def dev_rgb(torch, st, arr, pad=64):
    t = st.pitched(arr.shape[0], arr.shape[1], 3, pad)
    t.copy_(torch.from_numpy(np.ascontiguousarray(arr)).cuda())
    return t


def assert_within_one_lsb(out, ref, what):
    """The blend's bar: every 8-bit value within 1 of the oracle's. The GPU
    sums the level-1 Gaussian in another order (raw pull sums plus the filter
    of the fill), so a value whose exact blend sits at a rounding boundary can
    land on the other side; such pixels are rare, and a real error would also
    exceed 1 LSB."""
    diff = np.abs(out.astype(int) - ref.astype(int))
    assert diff.max() <= 1, (what, int(diff.max()))
    assert (diff > 0).mean() < 1e-3, (what, float((diff > 0).mean()))


def render_both(T, a, b, H, mode):
    torch, st = T
    status, g = st.prepare_geometry(a.shape[1], a.shape[0], b.shape[1], b.shape[0], np.asarray(H, np.float64))
    ost, og = O.prepare_geometry(a.shape[1], a.shape[0], b.shape[1], b.shape[0], H)
    assert status == ost == 0
    assert (g.width, g.height, g.origin_x, g.origin_y) == (og.width, og.height, og.origin_x, og.origin_y)
    assert list(g.inverse) == list(og.inverse)
    out, mask, counters = st.s6_render(dev_rgb(torch, st, a), dev_rgb(torch, st, b, pad=5), g, mode)
    torch.cuda.synchronize()
    oout, omask, ostats = O.s6_render(a, b, og, mode)
    return out.cpu().numpy(), mask.cpu().numpy(), tuple(int(v) for v in counters.cpu()), oout, omask, ostats


def test_s6_golden_vectors(T):
    gold = np.load(O.ROOT / "tests" / "golden" / "s6_golden.npz")
    n = 0
    for key in gold.files:
        if not key.endswith("__panorama"):
            continue
        base = key[:-10]
        mode = O.MODES[base.split("__")[1]]
        a, b, H = gold[base + "__a"], gold[base + "__b"], gold[base + "__h"]
        out, mask, _, oout, omask, _ = render_both(T, a, b, H, mode)
        assert np.array_equal(mask, gold[base + "__mask"]), base
        assert np.abs(out.astype(int) - gold[key].astype(int)).max() <= 1, base
        assert np.array_equal(mask, omask), base
        assert_within_one_lsb(out, oout, base)
        n += 1
    assert n == 33


@pytest.mark.parametrize("mode", [0, 2])
@pytest.mark.parametrize("snap", [0.0, 0.25])
def test_s6_render_snap_tolerances(T, mode, snap):
    """B's sample positions snapped under the reference's rule with other
    tolerances than the default: 0 (the threshold is the FP64 round-off term,
    the warp's general path) and 0.25 (every position within a quarter pixel
    of an integer snaps; the warp's short form, which holds when the
    tolerance exceeds 16 eps max(1, W, H)). A slight scale and a fractional
    shift put positions at every distance from the integers."""
    torch, st = T
    h, w = 120, 160
    a = images.texture_rgb(70 + mode, h, w)
    b = images.texture_rgb(90 + mode, h, w)
    H = np.array([[1.0003, 0.0, 0.45 * w + 0.125], [0.0, 0.9998, 0.07 * h - 0.375], [0.0, 0.0, 1.0]])
    pol = st.F.GeometryPolicy(1e-12, 1e-12, snap, 16777216)
    status, g = st.prepare_geometry(w, h, w, h, H, pol)
    ost, og = O.prepare_geometry(w, h, w, h, H, snap_tol=snap)
    assert status == ost == 0
    assert g.snap_tolerance == og.snap_tolerance == snap and list(g.inverse) == list(og.inverse)
    out, mask, counters = st.s6_render(dev_rgb(torch, st, a), dev_rgb(torch, st, b, pad=5), g, mode)
    torch.cuda.synchronize()
    oout, omask, ostats = O.s6_render(a, b, og, mode)
    assert np.array_equal(mask.cpu().numpy(), omask)
    assert_within_one_lsb(out.cpu().numpy(), oout, (mode, snap))
    assert tuple(int(v) for v in counters.cpu()) == ostats


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("size", [(120, 160), (480, 640), (1080, 1920)])
def test_s6_render_matches_oracle(T, mode, size):
    h, w = size
    a = images.texture_rgb(h + mode, h, w)
    th = 0.03
    H = np.array([[np.cos(th) * 1.02, -np.sin(th), 0.45 * w], [np.sin(th), np.cos(th) * 0.99, 0.07 * h],
                  [2e-5, -1e-5, 1.0]])
    b = images.texture_rgb(h + 100 + mode, h, w)
    out, mask, stats, oout, omask, ostats = render_both(T, a, b, H, mode)
    assert np.array_equal(mask, omask)
    assert_within_one_lsb(out, oout, (mode, size))
    assert stats == ostats


@pytest.mark.parametrize("mode", [0, 2, 4])
@pytest.mark.parametrize("th,persp", [(0.45, (3e-4, -2e-4)), (-0.8, (-4e-4, 1e-4)), (1.2, (0.0, 5e-4)),
                                      (np.pi / 2 + 0.05, (1e-4, 1e-4))])
def test_s6_render_steep_quads(T, mode, th, persp):
    """B far from axis-aligned (rotations of 26-95 degrees, perspective): its
    quad covers little of its bounding box, so the warp's per-row bounds from
    the quad's edges decide which pixels take the inverse map. Mask and
    counts exact, colour within 1 LSB of the oracle."""
    h, w = 240, 320
    a = images.texture_rgb(7 + mode, h, w)
    b = images.texture_rgb(70 + mode, h, w)
    H = np.array([[np.cos(th), -np.sin(th), 0.5 * w], [np.sin(th), np.cos(th), 0.2 * h], [persp[0], persp[1], 1.0]])
    out, mask, stats, oout, omask, ostats = render_both(T, a, b, H, mode)
    assert np.array_equal(mask, omask)
    assert stats == ostats
    assert_within_one_lsb(out, oout, (mode, th))
