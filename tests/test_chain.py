"""Chains of frames on one canvas (no counterpart in the reference, which
stitches pairs), on the CPU.

Pinning of the oracle's chain functions (orc_s6_prepare_chain,
orc_s6_render_chain): for two frames every expression is the pair render's,
so the chain render must equal orc_s6_render (itself pinned against the
reference's golden vectors) bit for bit in every blend mode, and the chain
canvas must equal the pair canvas; one frame renders to itself; n identical
frames under the identity give frame 0 back within rounding with the whole
canvas covered; pure integer translations give exactly the union box. The
product's chain geometry (cs_s6_prepare_chain) is compared with the oracle's
bit for bit, and the chain render's host argument rules are checked without
a GPU."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O

MODES = ["feather", "multiband_3", "multiband_5", "multiband_6", "multiband_7"]
INT_FIELDS = ["width", "height", "origin_x", "origin_y", "width_a", "height_a", "width_b", "height_b"]


def geometry_tuple(g):
    return tuple(getattr(g, f) for f in INT_FIELDS) + (g.denominator_tolerance, g.snap_tolerance) + tuple(g.inverse)


def same_bits(x, y):
    return np.array(x, np.float64).tobytes() == np.array(y, np.float64).tobytes()


def strip(seed, n, w, h, overlap=0.45, kind=0):
    from cs_stitch import pipeline
    return pipeline.synth_strip(seed, n, w, h, overlap, kind)


@pytest.mark.parametrize("seed,size,sizes_b", [(5, (96, 72), None), (6, (120, 80), (100, 70)), (7, (64, 64), None)])
def test_oracle_chain_of_two_is_the_pair_render(seed, size, sizes_b):
    from cs_stitch import pipeline
    w, h = size
    a, b, H = pipeline.synth_pair(seed, w, h, (0.35, 0.6))
    if sizes_b:
        b = np.ascontiguousarray(b[: sizes_b[1], : sizes_b[0]])
    wb, hb = b.shape[1], b.shape[0]
    st, g = O.prepare_chain([(w, h), (wb, hb)], [H])
    assert st == 0
    # The chain canvas is the pair canvas of the normalised map.
    Hn = H / np.abs(H).max()
    st2, gp = O.prepare_geometry(w, h, wb, hb, Hn)
    assert st2 == 0
    assert geometry_tuple(g[1]) == geometry_tuple(gp)
    for mode in MODES:
        m = O.MODES[mode]
        out_p, mask_p, stats_p = O.s6_render(a, b, g[1], m)
        out_c, mask_c, stats_c = O.s6_render_chain([a, b], g, m)
        assert np.array_equal(mask_p, mask_c), mode
        assert stats_p == stats_c, mode
        assert np.array_equal(out_p, out_c), (mode, int(np.abs(out_p.astype(int) - out_c).max()))


@pytest.mark.parametrize("mode", MODES)
def test_oracle_chain_of_one_is_the_frame(mode):
    frame = np.random.default_rng(3).integers(0, 256, (37, 53, 3), dtype=np.uint8)
    st, g = O.prepare_chain([(53, 37)], [])
    assert st == 0 and (g[0].width, g[0].height, g[0].origin_x, g[0].origin_y) == (53, 37, 0, 0)
    out, mask, stats = O.s6_render_chain([frame], g, mode)
    assert np.array_equal(out, frame)
    assert mask.all() and stats == (53 * 37, 0)


@pytest.mark.parametrize("mode", MODES)
def test_oracle_identical_frames_give_frame_zero(mode):
    # Every pixel is covered by all three frames with equal scores: frame 0
    # owns everything, the mix is frame 0's bands alone and the feather
    # average is its colour.
    frame = np.random.default_rng(4).integers(0, 256, (40, 56, 3), dtype=np.uint8)
    eye = np.eye(3)
    st, g = O.prepare_chain([(56, 40)] * 3, [eye, eye])
    assert st == 0 and (g[0].width, g[0].height) == (56, 40)
    out, mask, stats = O.s6_render_chain([frame, frame, frame], g, mode)
    assert mask.all() and stats == (56 * 40, 56 * 40)
    assert int(np.abs(out.astype(int) - frame).max()) <= 1


def test_oracle_integer_translations_give_the_union_box():
    hs = [np.array([[1, 0, 30], [0, 1, -7], [0, 0, 1]], np.float64),
          np.array([[1, 0, 25], [0, 1, 11], [0, 0, 1]], np.float64)]
    st, g = O.prepare_chain([(50, 40), (50, 40), (50, 40)], hs)
    assert st == 0
    # Frame 1 at (30, -7), frame 2 at (55, 4) in frame-0 pixels.
    assert (g[0].origin_x, g[0].origin_y, g[0].width, g[0].height) == (0, -7, 105, 51)
    frames = [np.full((40, 50, 3), v, np.uint8) for v in (10, 20, 30)]
    out, mask, stats = O.s6_render_chain(frames, g, "feather")
    # Canvas boxes: frame 0 x [0, 50) y [7, 47), frame 1 x [30, 80) y [0, 40),
    # frame 2 x [55, 105) y [11, 51).
    assert mask[11:51, 55:105].all() and not mask[0:7, 80:].any() and not mask[47:, :30].any()
    assert (out[45, 100] == 30).all() and (out[20, 10] == 10).all() and (out[2, 60] == 20).all()
    assert stats[0] == int(mask.sum())


def chain_cases():
    cases = []
    for seed, n, w, h in [(1, 3, 96, 64), (2, 5, 80, 60), (3, 8, 64, 48), (4, 2, 200, 150)]:
        _, hs = strip(seed, n, w, h)
        cases.append(([(w, h)] * n, hs, None))
    rng = np.random.default_rng(9)
    for _ in range(20):
        n = int(rng.integers(2, 7))
        sizes = [tuple(int(v) for v in rng.integers(8, 300, 2)) for _ in range(n)]
        hs = []
        for _ in range(n - 1):
            th, s = rng.uniform(-0.2, 0.2), rng.uniform(0.8, 1.25)
            hs.append(np.array([[s * np.cos(th), -s * np.sin(th), rng.uniform(-200, 200)],
                                [s * np.sin(th), s * np.cos(th), rng.uniform(-50, 50)],
                                [rng.uniform(-3e-4, 3e-4), rng.uniform(-3e-4, 3e-4), rng.uniform(0.5, 2)]]))
        cases.append((sizes, hs, None))
    eye = np.eye(3)
    cases.append(([(4000, 3000)] * 3, [np.array([[1, 0, 3000], [0, 1, 0], [0, 0, 1.0]])] * 2, 16777216))  # budget
    cases.append(([(64, 48)] * 3, [eye, np.zeros((3, 3))], None))                    # a zero step
    cases.append(([(64, 48)] * 3, [eye, np.array([[1, 2, 3], [2, 4, 6], [1, 2, 3.0]])], None))  # singular
    cases.append(([(64, 48), (0, 48)], [eye], None))                                  # invalid size
    cases.append(([(64, 48), (64, 48)], [np.array([[1, 0, np.nan], [0, 1, 0], [0, 0, 1.0]])], None))
    return cases


def test_product_chain_geometry_matches_oracle():
    from cs_stitch import _ffi as F
    from cs_stitch import stages
    statuses = set()
    for sizes, hs, budget in chain_cases():
        pol = F.GeometryPolicy(1e-12, 1e-12, 1e-10, budget if budget else 1 << 28)
        st, g = stages.prepare_chain(sizes, hs, pol)
        ost, og = O.prepare_chain(sizes, hs, max_pixels=budget if budget else 1 << 28)
        assert st == ost, (sizes, st, ost)
        statuses.add(st)
        if st != 0:
            continue
        for k in range(len(sizes)):
            pk, ok = geometry_tuple(g[k]), geometry_tuple(og[k])
            assert pk[:8] == ok[:8], (k, pk[:8], ok[:8])
            assert same_bits(pk[8:], ok[8:]), k
    assert statuses >= {0, 1, 2, 5}


def test_chain_canvas_holds_every_frame():
    from cs_stitch import stages
    for sizes, hs, _ in chain_cases()[:24]:
        st, g = stages.prepare_chain(sizes, hs)
        if st != 0:
            continue
        G = np.eye(3)
        for k in range(len(sizes)):
            if k:
                G = G @ hs[k - 1]
            w, h = sizes[k]
            for u, v in ((0, 0), (w - 1, 0), (0, h - 1), (w - 1, h - 1)):
                p = G @ np.array([u, v, 1.0])
                x, y = p[0] / p[2] - g[0].origin_x, p[1] / p[2] - g[0].origin_y
                assert -1e-6 <= x <= g[0].width - 1 + 1e-6 and -1e-6 <= y <= g[0].height - 1 + 1e-6, (k, x, y)
            # The inverse map takes the corner's canvas position back to it.
            inv = np.array(g[k].inverse).reshape(3, 3)
            p = G @ np.array([w - 1, h - 1, 1.0])
            q = inv @ (p / p[2])
            assert np.allclose(q[:2] / q[2], (w - 1, h - 1), atol=1e-6)


def test_synth_strip_frames_follow_their_homographies():
    frames, hs = strip(21, 4, 160, 120, overlap=0.5)
    again, hs2 = strip(21, 4, 160, 120, overlap=0.5)
    assert np.array_equal(frames, again) and all(np.array_equal(x, y) for x, y in zip(hs, hs2))
    for k, H in enumerate(hs):
        a, b = frames[k].astype(np.float64), frames[k + 1].astype(np.float64)
        diffs = []
        for y in range(10, 110, 4):
            for x in range(10, 150, 4):
                p = H @ np.array([x, y, 1.0])
                u, v = p[0] / p[2], p[1] / p[2]
                if 1 <= u < 158 and 1 <= v < 118:
                    diffs.append(b[y, x].mean() - a[int(round(v)), int(round(u))].mean())
        diffs = np.array(diffs)
        assert len(diffs) > 100
        # Same world up to the per-frame brightness offset, noise and resampling.
        assert np.median(np.abs(diffs - np.median(diffs))) < 12


OK, INVALID, INSUFFICIENT = 0, 1, 2


def test_chain_workspace_and_argument_rules():
    from cs_stitch import _ffi as F
    from test_s6_arguments import Arena
    L = F.load()
    _, hs = strip(2, 3, 48, 32)
    st, g = O.prepare_chain([(48, 32)] * 3, hs)
    assert st == 0
    garr = (F.Geometry * 3)(*[F.Geometry.from_buffer_copy(bytes(x)) for x in g])
    n = C.c_size_t()
    assert L.cs_s6_chain_workspace_bytes(3, garr, 0, C.byref(n)) == OK and n.value > 0
    feather = n.value
    assert L.cs_s6_chain_workspace_bytes(3, garr, 4, C.byref(n)) == OK and n.value > feather
    need = n.value
    assert L.cs_s6_chain_workspace_bytes(3, garr, 17, C.byref(n)) == INVALID
    assert L.cs_s6_chain_workspace_bytes(0, garr, 0, C.byref(n)) == INVALID
    assert L.cs_s6_chain_workspace_bytes(65, garr, 0, C.byref(n)) == INVALID
    # Frames that disagree on the canvas.
    bad = (F.Geometry * 3)(*[F.Geometry.from_buffer_copy(bytes(x)) for x in g])
    bad[2].width += 1
    assert L.cs_s6_chain_workspace_bytes(3, bad, 0, C.byref(n)) == INVALID
    bad = (F.Geometry * 3)(*[F.Geometry.from_buffer_copy(bytes(x)) for x in g])
    bad[1].inverse[0] = float("nan")
    assert L.cs_s6_chain_workspace_bytes(3, bad, 0, C.byref(n)) == INVALID
    # Render-time checks that enqueue nothing.
    W, H = g[0].width, g[0].height
    arena = Arena(need + 4 * (W * H * 3 + 4096) + (1 << 16))
    frames = (F.Image * 3)(*[F.Image(arena.take(48 * 32 * 3), 48, 32, 48 * 3, F.RGB8) for _ in range(3)])
    out = F.Image(arena.take(W * H * 3), W, H, W * 3, F.RGB8)
    mask = F.Image(arena.take(W * H), W, H, W, F.GRAY8)
    ws = arena.take(need + 256)

    def render(fr=frames, o=out, m=mask, w=ws, nb=need, mode=4, count=3):
        return L.cs_s6_render_chain(count, fr, garr, mode, C.byref(o), C.byref(m), C.c_void_p(w) if w else None,
                                    nb, None, None)
    assert render(nb=need - 1) == INSUFFICIENT
    assert render(w=None) == INSUFFICIENT
    assert render(w=ws + 1) == INVALID
    assert render(o=F.Image(out.data, W - 1, H, W * 3, F.RGB8)) == INVALID
    assert render(m=F.Image(mask.data, W, H, W - 1, F.GRAY8)) == INVALID
    assert render(m=F.Image(out.data, W, H, W, F.GRAY8)) == INVALID          # mask over the panorama
    assert render(w=frames[1].data // 256 * 256) == INVALID                 # workspace over a frame
    wrong = (F.Image * 3)(*[frames[0], frames[1], F.Image(frames[2].data, 47, 32, 48 * 3, F.RGB8)])
    assert render(fr=wrong) == INVALID
    over = (F.Image * 3)(*[frames[0], frames[1], F.Image(out.data, 48, 32, 48 * 3, F.RGB8)])
    assert render(fr=over) == INVALID                                       # a frame under the panorama
    assert L.cs_s6_render_chain(3, None, garr, 4, C.byref(out), C.byref(mask), C.c_void_p(ws), need, None,
                                None) == INVALID
