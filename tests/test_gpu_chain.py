"""The N-frame render (cs_s6_render_chain) on the GPU against the oracle's
full-canvas chain render (orc_s6_render_chain): a strip of 4 frames of
640x480 in 5 bands and the other blend modes, ragged frame sizes, a chain of
two against the pair render, a chain of one, 3 frames of 3840x2160 in 7
bands (the large strip's frame size and depth), and the longest chain (64
frames). The bar is the pair
render's: the mask and the union/overlap counts exact, every 8-bit value
within 1 of the oracle's."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    import torch
    from cs_stitch import stages
    assert torch.cuda.is_available()
    return torch, stages


def dev_rgb(torch, st, arr, pad=64):
    t = st.pitched(arr.shape[0], arr.shape[1], 3, pad)
    t.copy_(torch.from_numpy(np.ascontiguousarray(arr)).cuda())
    return t


def assert_close(out, ref, what):
    diff = np.abs(out.astype(int) - ref.astype(int))
    assert diff.max() <= 1, (what, int(diff.max()))
    assert (diff > 0).mean() < 1e-3, (what, float((diff > 0).mean()))


def render_both(T, frames, hs, mode):
    torch, st = T
    sizes = [(f.shape[1], f.shape[0]) for f in frames]
    status, g = st.prepare_chain(sizes, hs)
    assert status == 0
    print(f"chain n={len(frames)} canvas={g[0].width}x{g[0].height} mode={mode}", flush=True)
    dev = [dev_rgb(torch, st, f) for f in frames]
    out, mask, counters = st.s6_render_chain(dev, g, O.MODES[mode])
    torch.cuda.synchronize()
    ref_out, ref_mask, ref_stats = O.s6_render_chain(frames, g, O.MODES[mode])
    got_mask = mask.cpu().numpy()
    assert np.array_equal(got_mask, ref_mask), mode
    assert tuple(int(v) for v in counters.cpu()) == tuple(ref_stats), mode
    return out.cpu().numpy(), ref_out, g


def strip(seed, n, w, h, overlap=0.45):
    from cs_stitch import pipeline
    return pipeline.synth_strip(seed, n, w, h, overlap)


@pytest.mark.parametrize("mode", ["multiband_5", "feather", "multiband_3", "multiband_7"])
def test_strip_of_four_matches_oracle(T, mode):
    frames, hs = strip(40, 4, 640, 480)
    out, ref, g = render_both(T, list(frames), hs, mode)
    assert g[0].width > 1200
    assert_close(out, ref, mode)


def test_ragged_frames_match_oracle(T):
    frames, hs = strip(41, 3, 400, 300, overlap=0.5)
    # Cropping a frame at its origin keeps the homographies into it valid.
    ragged = [frames[0], np.ascontiguousarray(frames[1][:260, :370]), np.ascontiguousarray(frames[2][:300, :330])]
    for mode in ("multiband_5", "feather"):
        out, ref, _ = render_both(T, ragged, hs, mode)
        assert_close(out, ref, mode)


def test_chain_of_two_matches_pair_render(T):
    torch, st = T
    from cs_stitch import pipeline
    a, b, H = pipeline.synth_pair(42, 480, 360, (0.3, 0.5))
    status, g = st.prepare_chain([(480, 360), (480, 360)], [H])
    assert status == 0
    da, db = dev_rgb(torch, st, a), dev_rgb(torch, st, b)
    for mode in ("multiband_5", "multiband_6", "feather"):
        print(f"pair vs chain mode={mode}", flush=True)
        out_c, mask_c, cnt_c = st.s6_render_chain([da, db], g, O.MODES[mode])
        out_p, mask_p, cnt_p = st.s6_render(da, db, g[1], O.MODES[mode])
        torch.cuda.synchronize()
        assert torch.equal(mask_c, mask_p) and torch.equal(cnt_c, cnt_p), mode
        assert_close(out_c.cpu().numpy(), out_p.cpu().numpy(), mode)


def test_chain_of_one_is_the_frame(T):
    torch, st = T
    frame = np.random.default_rng(43).integers(0, 256, (90, 130, 3), dtype=np.uint8)
    status, g = st.prepare_chain([(130, 90)], [])
    assert status == 0
    for mode in ("multiband_5", "feather"):
        out, mask, counters = st.s6_render_chain([dev_rgb(torch, st, frame)], g, O.MODES[mode])
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), frame), mode
        assert bool(mask.all()) and tuple(int(v) for v in counters.cpu()) == (130 * 90, 0)


def test_full_size_frames_seven_bands(T):
    # The large-strip frame size: at 7 bands the per-frame windows reach
    # about a thousand pixels past each frame, and frames' windows overlap
    # deeply; the oracle renders the 9203x2233 canvas on full planes (tens of
    # seconds on the host).
    frames, hs = strip(60, 3, 3840, 2160, overlap=0.3)
    out, ref, g = render_both(T, list(frames), hs, "multiband_7")
    assert (g[0].width, g[0].height) == (9203, 2233)
    assert_close(out, ref, "multiband_7 3x3840x2160")


def test_longest_chain(T):
    # The most frames a chain may hold: the per-launch tables are largest and
    # most canvas pixels are covered by two frames.
    frames, hs = strip(70, 64, 96, 64, overlap=0.5)
    for mode in ("multiband_5", "feather"):
        out, ref, g = render_both(T, list(frames), hs, mode)
        assert (g[0].width, g[0].height) == (2792, 110)
        assert_close(out, ref, mode)
