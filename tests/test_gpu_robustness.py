"""Robustness checks of the reference's stage tests, on every stage:

- guard bytes: each workspace and each output array is allocated with a
  tail of 0xa5 bytes (and images with 0xa5 row padding) that no stage may
  touch, with capacities above the counts so that the last partial block of
  every kernel is exercised (tests/test_s4_cuda.cu:117-121, 176-180);
- no allocation and no leak: free device memory is the same before and
  after many repeated runs on the same buffers - the plain runner, the
  batched form and the served form (tests/test_s2_s3_cuda.cu:160-171).
"""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GUARD = 256
FILL = 0xA5


def guarded(nbytes):
    """uint8 device buffer of nbytes plus a guard tail, all FILL."""
    return torch.full((max(nbytes, 0) + GUARD,), FILL, dtype=torch.uint8, device="cuda")


def tail_intact(buf, nbytes):
    return bool((buf[nbytes:] == FILL).all().item())


def padded_image(h, w, channels, pad=64):
    """Pitched image whose row padding holds FILL; returns (view, base)."""
    base = torch.full((h, w * channels + pad), FILL, dtype=torch.uint8, device="cuda")
    v = base[:, : w * channels]
    return (v.view(h, w, channels) if channels == 3 else v), base


def padding_intact(base, w, channels):
    return bool((base[:, w * channels:] == FILL).all().item())


def typed(buf, n, dtype):
    size = torch.tensor([], dtype=dtype).element_size()
    return buf[: n * size].view(dtype)


@pytest.mark.parametrize("w,h,K,R,blend", [(1920, 1080, 5000, 4096, "multiband_6"), (640, 480, 1500, 512, "feather"),
                                           (301, 203, 700, 256, "multiband_3")])
def test_every_stage_stays_inside_its_buffers(w, h, K, R, blend):
    from cs_stitch import _ffi as F
    from cs_stitch import pipeline as P
    from cs_stitch import stages as st
    L = F.load()
    stream = st.stream_handle()
    a, b, _ = P.synth_pair(41000 + w, w, h, overlap=(0.5, 0.7))
    ra, ra_base = padded_image(h, w, 3)
    rb, rb_base = padded_image(h, w, 3)
    ra.copy_(torch.from_numpy(a).cuda())
    rb.copy_(torch.from_numpy(b).cuda())
    # S1
    ga, ga_base = padded_image(h, w, 1)
    gb, gb_base = padded_image(h, w, 1)
    sa, sb = F.ScaleInfo(), F.ScaleInfo()
    ia, ib, oa, ob = st.image(ra), st.image(rb), st.image(ga), st.image(gb)
    assert L.cs_s1_preprocess(C.byref(ia), C.byref(ib), 1.0, C.byref(oa), C.byref(ob), C.byref(sa), C.byref(sb),
                              stream) == 0
    # S2 with output capacity K + 37 and a workspace with a guard tail
    cap = K + 37
    kps = []
    for g in (ga, gb):
        bufs = [guarded(4 * cap) for _ in range(3)]
        count = guarded(4)
        kp = st.KeypointSet(typed(bufs[0], cap, torch.float32), typed(bufs[1], cap, torch.float32),
                            typed(bufs[2], cap, torch.float32), typed(count, 1, torch.int32))
        n = C.c_size_t()
        assert L.cs_s2_workspace_bytes(w, h, C.byref(F.S2Config(20, K, 0.04)), C.byref(n)) == 0
        ws = guarded(n.value)
        st.s2_detect(g, 20, K, out=kp, ws=ws[: n.value])
        kps.append((kp, bufs, count, ws, n.value))
    # S3
    descs = []
    n3 = C.c_size_t()
    assert L.cs_s3_workspace_bytes(w, h, C.byref(n3)) == 0
    for g, (kp, *_rest) in zip((ga, gb), kps):
        ws3 = guarded(n3.value)
        dbuf = guarded(32 * cap)
        d = F.Descriptors(F.BINARY_U8, 32, 32, dbuf.data_ptr())
        img, k = st.image(g), kp.c(cap)
        assert L.cs_s3_describe(C.byref(img), C.byref(k), C.byref(d), ws3.data_ptr(), n3.value, stream) == 0
        descs.append((dbuf, ws3))
    # S4 (cross-check), match arrays of capacity cap
    cfg4 = F.S4Config(0.8, 1)
    n4 = C.c_size_t()
    assert L.cs_s4_workspace_bytes(cap, cap, C.byref(cfg4), C.byref(n4)) == 0
    ws4 = guarded(n4.value)
    mb = [guarded(4 * cap) for _ in range(3)]
    mcount = guarded(4)
    m = F.Matches(mb[0].data_ptr(), mb[1].data_ptr(), mb[2].data_ptr(), cap, mcount.data_ptr())
    ka, kb = kps[0][0].c(cap), kps[1][0].c(cap)
    ea = F.Descriptors(F.BINARY_U8, 32, 32, descs[0][0].data_ptr())
    eb = F.Descriptors(F.BINARY_U8, 32, 32, descs[1][0].data_ptr())
    assert L.cs_s4_match(C.byref(ka), C.byref(ea), C.byref(kb), C.byref(eb), C.byref(cfg4), C.byref(m),
                         ws4.data_ptr(), n4.value, stream) == 0
    # S5
    cfg5 = F.S5Config(R, 0, 64, 1, 3.0, 0, 1)
    n5 = C.c_size_t()
    assert L.cs_s5_workspace_bytes(C.byref(cfg5), C.byref(n5)) == 0
    ws5 = guarded(n5.value)
    mask5 = guarded(cap)
    pinned = torch.zeros(C.sizeof(F.HomographyResult), dtype=torch.uint8, pin_memory=True)
    assert L.cs_s5_estimate(C.byref(ka), C.byref(kb), C.byref(m), C.byref(sa), C.byref(sb), C.byref(cfg5),
                            mask5.data_ptr(), pinned.data_ptr(), ws5.data_ptr(), n5.value, stream) == 0
    torch.cuda.synchronize()
    res = F.HomographyResult.from_buffer_copy(pinned.numpy().tobytes())
    assert res.status == 0, res.status
    # S6
    status, geo = st.prepare_geometry(w, h, w, h, np.array(res.H[:]))
    assert status == 0
    mode = F.BLEND_NAMES[blend]
    n6 = C.c_size_t()
    assert L.cs_s6_workspace_bytes(C.byref(geo), mode, C.byref(n6)) == 0
    ws6 = guarded(n6.value)
    pano, pano_base = padded_image(geo.height, geo.width, 3)
    pmask, pmask_base = padded_image(geo.height, geo.width, 1)
    counters = guarded(16)
    io, im = st.image(pano), st.image(pmask)
    assert L.cs_s6_render(C.byref(ia), C.byref(ib), C.byref(geo), mode, C.byref(io), C.byref(im),
                          ws6.data_ptr() if n6.value else None, n6.value, stream,
                          C.cast(counters.data_ptr(), C.POINTER(F.RenderStatistics))) == 0
    torch.cuda.synchronize()
    # Every guard is intact.
    for base, cw, ch in ((ra_base, w, 3), (rb_base, w, 3), (ga_base, w, 1), (gb_base, w, 1),
                         (pano_base, geo.width, 3), (pmask_base, geo.width, 1)):
        assert padding_intact(base, cw, ch)
    for kp, bufs, count, ws, nbytes in kps:
        assert all(tail_intact(x, 4 * cap) for x in bufs) and tail_intact(count, 4) and tail_intact(ws, nbytes)
        assert 0 < int(kp.count.item()) <= K
    for dbuf, ws3 in descs:
        assert tail_intact(dbuf, 32 * cap) and tail_intact(ws3, n3.value)
    assert all(tail_intact(x, 4 * cap) for x in mb) and tail_intact(mcount, 4) and tail_intact(ws4, n4.value)
    assert tail_intact(ws5, n5.value) and tail_intact(mask5, cap)
    assert tail_intact(ws6, n6.value) and tail_intact(counters, 16)


def free_bytes():
    torch.cuda.synchronize()
    return torch.cuda.mem_get_info()[0]


def small_pairs(n, seed0=42000):
    from cs_stitch import pipeline as P
    from cs_stitch import stages as st
    out = []
    for i in range(n):
        w, h = [(640, 480), (800, 600), (320, 240)][i % 3]
        a, b, _ = P.synth_pair(seed0 + i, w, h, overlap=(0.5, 0.8))
        ta, tb = st.pitched(h, w, 3), st.pitched(h, w, 3)
        ta.copy_(torch.from_numpy(a).cuda())
        tb.copy_(torch.from_numpy(b).cuda())
        out.append((ta, tb))
    torch.cuda.synchronize()
    return out


def test_repeated_runs_allocate_nothing():
    from cs_stitch import batch as B
    from cs_stitch import pipeline as P
    pairs = small_pairs(6)
    # Plain runner: 1000 runs on the same buffers.
    runner = P.PipelineRunner(800, 600, max_k=1000, max_r=512, canvas_width=2048, canvas_height=1024,
                              workspace_mode=P.F.BLEND_NAMES["multiband_5"])
    cfg = P.harness_config(S=1.0, K=800, R=256, blend="multiband_5", cross_check=True)
    runner.use_device_pair(*pairs[0])
    first = runner.execute_run(cfg)
    before = free_bytes()
    for i in range(1000):
        runner.use_device_pair(*pairs[i % len(pairs)])
        rec = runner.execute_run(cfg)
    assert free_bytes() == before
    runner.use_device_pair(*pairs[0])
    again = runner.execute_run(cfg)
    assert (again.status, list(again.H), again.num_inliers) == (first.status, list(first.H), first.num_inliers)
    runner.close()
    # Batched form: 1000 batches of 6 pairs, graph replays on the pool.
    br = B.BatchRunner(max_pairs=6, pool_bytes=2 << 30)
    specs = [B.pair_spec(ta, tb, K=800, R=256, blend="multiband_5", cross_check=True) for ta, tb in pairs]
    ref, _ = br.run(specs)
    br.run(specs)  # both pool slots now hold a batch (outputs stay valid until the second run after)
    _, used0, _ = br.pool()
    before = free_bytes()
    for _ in range(1000):
        outs, _ = br.run(specs)
    assert free_bytes() == before
    _, used1, _ = br.pool()
    assert used1 == used0
    assert [o.status for o in outs] == [o.status for o in ref]
    assert [list(o.homography.H) for o in outs] == [list(o.homography.H) for o in ref]
    br.close()
    # Served form: two traces of 300 requests leave the pool and device memory as they were.
    srv = B.Server(hbm_budget=2 << 30, max_batch=16)
    trace = [B.request(i, *pairs[i % len(pairs)], 0.05 * i, seed=1) for i in range(300)]
    done1, _ = srv.run_trace(trace)
    before = free_bytes()
    done2, st2 = srv.run_trace(trace)
    assert free_bytes() == before
    assert sorted((c.id, c.status, list(c.homography.H)) for c in done1) == \
        sorted((c.id, c.status, list(c.homography.H)) for c in done2)
    srv.close()
