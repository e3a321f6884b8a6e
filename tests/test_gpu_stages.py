"""GPU parity of S1-S4 against the CPU oracle through the C ABI.

Bar: bit-exact gray images, keypoints (x, y, response), descriptors, match
indices and distances.
"""
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


def to_dev(torch, arr):
    return torch.from_numpy(np.ascontiguousarray(arr)).cuda()


def rgb_dev(torch, stages, rgb, pad=64):
    h, w, _ = rgb.shape
    t = stages.pitched(h, w, 3, pad)
    t.copy_(torch.from_numpy(rgb).cuda())
    return t


@pytest.mark.parametrize("shape", [(480, 640), (67, 93), (1080, 1920), (5, 3)])
@pytest.mark.parametrize("scale", [1.0, 0.5, 0.25, 0.75, 0.6])
def test_s1_preprocess_matches_oracle(T, shape, scale):
    torch, st = T
    h, w = shape
    rng = np.random.default_rng(h * 7 + w)
    a = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    b = images.texture_rgb(w, h, w) if min(h, w) > 8 else rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    ga, gb, sa, sb = st.s1_preprocess(rgb_dev(torch, st, a), rgb_dev(torch, st, b, pad=13), scale)
    torch.cuda.synchronize()
    assert np.array_equal(ga.cpu().numpy(), O.s1_preprocess(a, scale))
    assert np.array_equal(gb.cpu().numpy(), O.s1_preprocess(b, scale))
    _, si = O.scale_info(w, h, scale)
    assert (sa.work_w, sa.work_h) == (si.work_w, si.work_h) and list(sa.D) == list(si.D)


def test_s1_gradient_known_answers(T):
    torch, st = T
    const = torch.full((40, 50), 77, dtype=torch.uint8, device="cuda")
    yy, xx = np.mgrid[0:30, 0:20]
    checker = ((xx + yy) % 2 * 255).astype(np.uint8)
    ga, gb = st.s1_compute_gradient(const, to_dev(torch, checker))
    assert ga == 0.0 and gb == 1.0
    rng = np.random.default_rng(3)
    g = rng.integers(0, 256, (57, 91), dtype=np.uint8)
    x, _ = st.s1_compute_gradient(to_dev(torch, g), to_dev(torch, g))
    assert x == O.s1_gradient(g)


def detect_and_describe(torch, st, gray_np, thr, k):
    g = to_dev(torch, gray_np)
    kp = st.s2_detect(g, thr, k)
    desc = st.s3_describe(g, kp, capacity=k)
    torch.cuda.synchronize()
    n, x, y, r = kp.host()
    return np.stack([x.numpy(), y.numpy(), r.numpy()], axis=1) if n else np.zeros((0, 3), np.float32), \
        desc[:n].cpu().numpy()


def test_s2_s3_golden_vectors(T):
    """The reference's NumPy ORB outputs (tests/golden) reproduced on the GPU."""
    torch, st = T
    gold = np.load(O.ROOT / "tests" / "golden" / "s2s3_golden.npz")
    n = 0
    for key in gold.files:
        if not key.endswith("__kp"):
            continue
        base = key[:-4]
        name = base.split("__")[0]
        k = int(base.split("__k")[1].split("__")[0])
        thr = int(base.split("__t")[1])
        if k == 0:
            continue
        kp, desc = detect_and_describe(torch, st, gold[name + "__gray"], thr, k)
        assert kp.shape == gold[key].shape, base
        assert np.array_equal(kp, gold[key]), base
        assert np.array_equal(desc, gold[base + "__desc"]), base
        n += 1
    assert n >= 40


# The top-K order sorts 2048-key segments and merges them by binary search:
# K 1000 (one segment of 2000 candidates), 5000 (5 segments), 8000 and 16000
# (8 and 16 segments, the large strip's keypoint count).
@pytest.mark.parametrize("seed,h,w,k", [(1, 480, 640, 1000), (2, 720, 1280, 5000), (3, 1080, 1920, 5000),
                                        (4, 300, 451, 37), (5, 2160, 3840, 8000), (6, 2160, 3840, 16000),
                                        (7, 1080, 1920, 1024)])
def test_s2_s3_match_oracle(T, seed, h, w, k):
    torch, st = T
    gray = images.gray_of(images.texture_rgb(seed, h, w, blobs=200, rects=50))
    kp, desc = detect_and_describe(torch, st, gray, 20, k)
    ref = O.s2_detect(gray, 20, k)
    assert kp.shape == ref.shape
    assert np.array_equal(kp, ref)
    assert np.array_equal(desc, O.s3_describe(gray, ref))


def test_s2_zero_k_and_tiny_image(T):
    torch, st = T
    g = torch.randint(0, 256, (62, 200), dtype=torch.uint8, device="cuda")
    kp = st.s2_detect(g, 20, 100)
    torch.cuda.synchronize()
    assert int(kp.count.item()) == 0
    g2 = torch.randint(0, 256, (100, 100), dtype=torch.uint8, device="cuda")
    kp2 = st.s2_detect(g2, 20, 0)
    torch.cuda.synchronize()
    assert int(kp2.count.item()) == 0


def planted_descriptors(seed, na, nb, dup=0.5):
    rng = np.random.default_rng(seed)
    da = rng.integers(0, 256, (na, 32), dtype=np.uint8)
    db = rng.integers(0, 256, (nb, 32), dtype=np.uint8)
    # Near-duplicates: B rows copied from A with a few flipped bits, in shuffled order.
    m = int(min(na, nb) * dup)
    src = rng.permutation(na)[:m]
    dst = rng.permutation(nb)[:m]
    for s, d in zip(src, dst):
        row = da[s].copy()
        for _ in range(rng.integers(0, 12)):
            row[rng.integers(0, 32)] ^= np.uint8(1 << rng.integers(0, 8))
        db[d] = row
    # Exact ties: duplicate some B rows so equal distances must go to the lower index.
    for _ in range(max(1, nb // 50)):
        i, j = rng.integers(0, nb, 2)
        db[j] = db[i]
    return da, db


@pytest.mark.parametrize("na,nb", [(0, 10), (1, 1), (2, 3), (100, 80), (1000, 1000), (5000, 5000), (777, 3001),
                                   (3001, 777)])
@pytest.mark.parametrize("cross", [False, True])
def test_s4_match_matches_oracle(T, na, nb, cross):
    torch, st = T
    da, db = planted_descriptors(na * 31 + nb, na, nb)
    ca = torch.tensor([na], dtype=torch.int32, device="cuda")
    cb = torch.tensor([nb], dtype=torch.int32, device="cuda")
    # Capacities above the counts: rows past the count must be ignored.
    pad_a = np.vstack([da, np.full((13, 32), 0xAB, np.uint8)])
    pad_b = np.vstack([db, np.full((7, 32), 0xCD, np.uint8)])
    out, _ = st.s4_match(to_dev(torch, pad_a), ca, to_dev(torch, pad_b), cb, 0.8, cross)
    torch.cuda.synchronize()
    ia, ib, dist = out.host()
    ra, rb, rd = O.s4_match(da, db, 0.8, cross)
    assert np.array_equal(ia.numpy(), ra) and np.array_equal(ib.numpy(), rb) and np.array_equal(dist.numpy(), rd)


@pytest.mark.parametrize("segment", [64, 256, 512])
def test_s4_result_independent_of_segment(T, segment):
    torch, st = T
    da, db = planted_descriptors(11, 1500, 2100)
    ca = torch.tensor([1500], dtype=torch.int32, device="cuda")
    cb = torch.tensor([2100], dtype=torch.int32, device="cuda")
    L = st.F.load()
    try:
        assert L.cs_s4_set_segment(segment) == 0
        out, _ = st.s4_match(to_dev(torch, da), ca, to_dev(torch, db), cb, 0.8, True)
        torch.cuda.synchronize()
        ia, ib, dist = out.host()
    finally:
        L.cs_s4_set_segment(128)
    ra, rb, rd = O.s4_match(da, db, 0.8, True)
    assert np.array_equal(ia.numpy(), ra) and np.array_equal(ib.numpy(), rb) and np.array_equal(dist.numpy(), rd)


def test_s4_forward_top2_and_edge_rules(T):
    """Edge answers of tests/test_s4_cuda.cu:185-215: ties to the lower index,
    a single train row has no second neighbour, zero distances need d2 > 0."""
    torch, st = T
    q = np.zeros((3, 32), np.uint8)
    q[1, 0] = 0b111
    t = np.zeros((4, 32), np.uint8)
    t[0, 0] = 0b1
    t[1, 0] = 0b10  # same distance to q0 as t0: tie -> lower index first
    t[2] = q[1]
    t[3] = q[1]     # exact duplicate: d1 == d2 == 0 never passes
    ca = torch.tensor([3], dtype=torch.int32, device="cuda")
    cb = torch.tensor([4], dtype=torch.int32, device="cuda")
    rec = st.top2_records(st.s4_knn_top2(to_dev(torch, q), ca, to_dev(torch, t), cb), 3)
    assert rec == [tuple(r) for r in O.s4_top2(q, t)]
    assert rec[0][2:] == (0, 1) and rec[1][:3] == (0.0, 0.0, 2)
    out, ws = st.s4_match(to_dev(torch, q), ca, to_dev(torch, t), cb, 0.8, False)
    torch.cuda.synchronize()
    ia, ib, _ = out.host()
    ra, rb, _ = O.s4_match(q, t, 0.8, False)
    assert np.array_equal(ia.numpy(), ra) and np.array_equal(ib.numpy(), rb)
    one = torch.tensor([1], dtype=torch.int32, device="cuda")
    rec1 = st.top2_records(st.s4_knn_top2(to_dev(torch, q), ca, to_dev(torch, t[:1]), one), 3)
    assert all(r[3] == -1 and r[1] == float("inf") for r in rec1)


def test_s4_invalid_counts_give_no_matches(T):
    torch, st = T
    da, db = planted_descriptors(5, 50, 50)
    bad = torch.tensor([51], dtype=torch.int32, device="cuda")  # count above capacity
    good = torch.tensor([50], dtype=torch.int32, device="cuda")
    out, _ = st.s4_match(to_dev(torch, da), bad, to_dev(torch, db), good, 0.8, True)
    torch.cuda.synchronize()
    assert int(out.count.item()) == 0


def test_s4_float32_l2(T):
    torch, st = T
    rng = np.random.default_rng(9)
    da = rng.normal(size=(300, 24)).astype(np.float32)
    db = rng.normal(size=(250, 24)).astype(np.float32)
    db[:100] = da[:100] + rng.normal(scale=0.05, size=(100, 24)).astype(np.float32)
    ca = torch.tensor([300], dtype=torch.int32, device="cuda")
    cb = torch.tensor([250], dtype=torch.int32, device="cuda")
    for cross in (False, True):
        out, _ = st.s4_match(to_dev(torch, da), ca, to_dev(torch, db), cb, 0.8, cross)
        torch.cuda.synchronize()
        ia, ib, dist = out.host()
        ra, rb, rd = O.s4_match(da, db, 0.8, cross)
        assert np.array_equal(ia.numpy(), ra) and np.array_equal(ib.numpy(), rb)
        assert np.array_equal(dist.numpy(), rd)
