"""cs_s4_match_noalloc (S4 matching in a fixed device array, no allocation)
through the C ABI against cs_s4_match with a caller workspace: the same
matches bit for bit on two calls in a row, the second with smaller inputs
(so stale contents of the shared array from the first call must not leak
into the second), with and without cross-check, plus the refusal of
capacities beyond the array."""
import ctypes as C

import numpy as np
import pytest
import torch

from test_gpu_stages import planted_descriptors

pytestmark = pytest.mark.gpu


def c_inputs(da, db, pad_a=11, pad_b=5):
    """Device copies with capacities above the counts, and their C views."""
    from cs_stitch import _ffi as F
    from cs_stitch import stages as st
    ta = torch.from_numpy(np.vstack([da, np.full((pad_a, 32), 0xAB, np.uint8)])).cuda()
    tb = torch.from_numpy(np.vstack([db, np.full((pad_b, 32), 0xCD, np.uint8)])).cuda()
    ca = torch.tensor([len(da)], dtype=torch.int32, device="cuda")
    cb = torch.tensor([len(db)], dtype=torch.int32, device="cuda")
    keep = (ta, tb, ca, cb)
    ka = F.Keypoints(None, None, None, ta.shape[0], ca.data_ptr())
    kb = F.Keypoints(None, None, None, tb.shape[0], cb.data_ptr())
    return keep, ka, st.descriptors(ta), kb, st.descriptors(tb)


def match(L, inputs, cross, noalloc):
    from cs_stitch import _ffi as F
    from cs_stitch import stages as st
    _, ka, ea, kb, eb = inputs
    cfg = F.S4Config(0.8, int(cross))
    out = st.MatchSet.empty(ka.capacity)
    m = out.c()
    stream = st.stream_handle()
    if noalloc:
        status = L.cs_s4_match_noalloc(C.byref(ka), C.byref(ea), C.byref(kb), C.byref(eb), C.byref(cfg),
                                       C.byref(m), stream)
    else:
        n = C.c_size_t()
        assert L.cs_s4_workspace_bytes(ka.capacity, kb.capacity, C.byref(cfg), C.byref(n)) == 0
        ws = st.workspace(n.value)
        status = L.cs_s4_match(C.byref(ka), C.byref(ea), C.byref(kb), C.byref(eb), C.byref(cfg), C.byref(m),
                               ws.data_ptr(), n.value, stream)
    assert status == 0, status
    torch.cuda.synchronize()
    ia, ib, d = out.host()
    return ia.numpy(), ib.numpy(), d.numpy()


@pytest.mark.parametrize("cross", [False, True])
def test_noalloc_matches_workspace_entry_twice_then_smaller(cross):
    from cs_stitch import _ffi as F
    L = F.load()
    for na, nb in ((9000, 8500), (1200, 700)):
        print("noalloc", cross, na, nb, flush=True)
        inputs = c_inputs(*planted_descriptors(na * 7 + nb, na, nb))
        want = match(L, inputs, cross, noalloc=False)
        got = match(L, inputs, cross, noalloc=True)
        assert len(want[0]) > 0.2 * min(na, nb)
        for w, g in zip(want, got):
            assert np.array_equal(w, g)


def test_noalloc_refuses_capacities_beyond_its_array():
    from cs_stitch import _ffi as F
    L = F.load()
    cfg = F.S4Config(0.8, 1)
    n = C.c_size_t()
    assert L.cs_s4_workspace_bytes(16384, 16384, C.byref(cfg), C.byref(n)) == 0
    assert n.value <= F.S4_NOALLOC_SCRATCH_BYTES
    assert L.cs_s4_workspace_bytes(20000, 20000, C.byref(cfg), C.byref(n)) == 0
    assert n.value > F.S4_NOALLOC_SCRATCH_BYTES
    inputs = c_inputs(*planted_descriptors(3, 20000, 20000, dup=0.1), pad_a=0, pad_b=0)
    _, ka, ea, kb, eb = inputs
    from cs_stitch import stages as st
    out = st.MatchSet.empty(ka.capacity)
    m = out.c()
    status = L.cs_s4_match_noalloc(C.byref(ka), C.byref(ea), C.byref(kb), C.byref(eb), C.byref(cfg), C.byref(m),
                                   st.stream_handle())
    assert status == 2  # INSUFFICIENT_WORKSPACE, nothing enqueued
