"""S6 host-side argument rules (the reference's tests/test_s6_cuda.cu
test_validation and the rejections of test_statistics): every case here is
refused before anything is enqueued, so the addresses are never read and
the checks run on the CPU. Workspace size and alignment, views that are null
or whose pitch is shorter than a row, and every forbidden overlap of a
written buffer (panorama, mask, statistics, workspace) with another one."""
import ctypes as C

import numpy as np
import pytest

OK, INVALID, INSUFFICIENT = 0, 1, 2


@pytest.fixture(scope="module")
def L():
    from cs_stitch import _ffi as F
    return F.load()


def geometry(F, w, h, wa, ha, wb, hb, inverse=(1, 0, 0, 0, 1, 0, 0, 0, 1)):
    return F.Geometry(w, h, 0, 0, wa, ha, wb, hb, 1e-12, 1e-10, (C.c_double * 9)(*inverse))


class Arena:
    """Host memory standing in for device addresses (only compared, never read)."""

    def __init__(self, nbytes):
        self.buf = np.zeros(nbytes + 4096, np.uint8)
        self.base = (self.buf.ctypes.data + 4095) // 4096 * 4096
        self.cursor = 0

    def take(self, nbytes):
        at = self.base + self.cursor
        self.cursor += (nbytes + 255) // 256 * 256
        return at


def views(F, arena, w=5, h=3):
    def image(channels):
        pitch = w * channels
        return F.Image(arena.take(pitch * h), w, h, pitch, F.RGB8 if channels == 3 else F.GRAY8)
    return image(3), image(3), image(3), image(1)


def render(L, a, b, g, mode, out, mask, ws, nbytes, stats=None):
    from cs_stitch import _ffi as F
    counters = C.cast(C.c_void_p(stats), C.POINTER(F.RenderStatistics)) if stats else None
    return L.cs_s6_render(C.byref(a), C.byref(b), C.byref(g), mode, C.byref(out), C.byref(mask),
                          C.c_void_p(ws) if ws else None, nbytes, None, counters)


def test_workspace_query_rules(L):
    from cs_stitch import _ffi as F
    g = geometry(F, 5, 3, 5, 3, 5, 3)
    n = C.c_size_t(123)
    assert L.cs_s6_workspace_bytes(C.byref(g), F.BLEND_NAMES["feather"], C.byref(n)) == OK and n.value == 0
    assert L.cs_s6_workspace_bytes(C.byref(g), 17, C.byref(n)) == INVALID
    assert L.cs_s6_workspace_bytes(C.byref(g), F.BLEND_NAMES["feather"], None) == INVALID
    broken = geometry(F, 5, 3, 5, 3, 5, 3, (float("nan"), 0, 0, 0, 1, 0, 0, 0, 1))
    assert L.cs_s6_workspace_bytes(C.byref(broken), F.BLEND_NAMES["feather"], C.byref(n)) == INVALID
    # A canvas that does not hold A, a zero map.
    assert L.cs_s6_workspace_bytes(C.byref(geometry(F, 4, 3, 5, 3, 5, 3)), 0, C.byref(n)) == INVALID
    assert L.cs_s6_workspace_bytes(C.byref(geometry(F, 5, 3, 5, 3, 5, 3, (0,) * 9)), 0, C.byref(n)) == INVALID


@pytest.mark.parametrize("blend", ["multiband_3", "multiband_5", "multiband_6", "multiband_7"])
def test_workspace_size_and_alignment(L, blend):
    from cs_stitch import _ffi as F
    mode = F.BLEND_NAMES[blend]
    g = geometry(F, 5, 3, 5, 3, 5, 3)
    n = C.c_size_t()
    assert L.cs_s6_workspace_bytes(C.byref(g), mode, C.byref(n)) == OK and n.value > 0
    arena = Arena(n.value + 4096)
    a, b, out, mask = views(F, arena)
    ws = arena.take(n.value + 256)
    assert render(L, a, b, g, mode, out, mask, ws, n.value - 1) == INSUFFICIENT
    assert render(L, a, b, g, mode, out, mask, None, n.value) == INSUFFICIENT
    assert render(L, a, b, g, mode, out, mask, ws + 1, n.value) == INVALID
    # The workspace may not cover an image.
    assert render(L, a, b, g, mode, out, mask, a.data, n.value) == INVALID
    assert render(L, a, b, g, mode, out, mask, mask.data - 256, n.value) == INVALID


def test_views_and_aliasing(L):
    from cs_stitch import _ffi as F
    g = geometry(F, 5, 3, 5, 3, 5, 3)
    arena = Arena(1 << 14)
    a, b, out, mask = views(F, arena)
    feather = F.BLEND_NAMES["feather"]
    assert render(L, a, b, g, feather, a, mask, None, 0) == INVALID       # panorama over an input
    bad = F.Image(out.data, 5, 3, 14, F.RGB8)                             # pitch shorter than a row
    assert render(L, a, b, g, feather, bad, mask, None, 0) == INVALID
    bad = F.Image(0, 5, 3, 15, F.RGB8)
    assert render(L, a, b, g, feather, bad, mask, None, 0) == INVALID
    bad = F.Image(out.data, 5, 3, 5, F.GRAY8)                             # mask over the panorama
    assert render(L, a, b, g, feather, out, bad, None, 0) == INVALID
    wrong = F.Image(out.data, 5, 3, 15, F.GRAY8)                          # format
    assert render(L, a, b, g, feather, wrong, mask, None, 0) == INVALID
    small = F.Image(out.data, 4, 3, 15, F.RGB8)                           # size differs from the canvas
    assert render(L, a, b, g, feather, small, mask, None, 0) == INVALID


@pytest.mark.parametrize("blend", ["feather", "multiband_3", "multiband_6"])
def test_statistics_aliasing(L, blend):
    from cs_stitch import _ffi as F
    mode = F.BLEND_NAMES[blend]
    g = geometry(F, 5, 3, 5, 3, 5, 3)
    n = C.c_size_t()
    assert L.cs_s6_workspace_bytes(C.byref(g), mode, C.byref(n)) == OK
    arena = Arena(n.value + 8192)
    a, b, out, mask = views(F, arena)
    ws = arena.take(max(n.value, 1))
    counters = arena.take(16)
    for bad in (a.data, out.data, mask.data, counters + 1):
        assert render(L, a, b, g, mode, out, mask, ws, n.value, bad) == INVALID, bad
    if n.value:
        assert render(L, a, b, g, mode, out, mask, ws, n.value, ws) == INVALID
        assert render(L, a, b, g, mode, out, mask, ws, n.value - 1, counters) == INSUFFICIENT
    else:
        # A workspace the mode does not need still may not hold the counters.
        assert render(L, a, b, g, mode, out, mask, ws, 256, ws) == INVALID
    bad = F.Image(0, 5, 3, 15, F.RGB8)
    assert render(L, a, b, g, mode, bad, mask, ws, n.value, counters) == INVALID
