"""The product's S6 canvas geometry (cs_s6_prepare_geometry, host long
double) against the reference's src/s6/geometry.cpp compiled from its
sources (oracle/_ref) and against the oracle's restatement: statuses, canvas
boxes and the inverse map bit for bit, over ordinary pairs and the edge cases
the contract names (src/s6/geometry.hpp:30-38): mixed denominator signs
across B's corners, near-singular maps, canvases over the pixel budget,
integer corners that must snap, and invalid arguments."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from test_oracle import dbl, ref_lib


def product(wa, ha, wb, hb, h, max_pixels=None):
    from cs_stitch import _ffi as F
    from cs_stitch import stages as st
    pol = None
    if max_pixels is not None:
        pol = F.GeometryPolicy(1e-12, 1e-12, 1e-10, max_pixels)
    return st.prepare_geometry(wa, ha, wb, hb, np.asarray(h, np.float64), pol)


def cases():
    rng = np.random.default_rng(11)
    out = []
    for _ in range(400):
        th = rng.uniform(-0.2, 0.2)
        s = rng.uniform(0.8, 1.25)
        h = [s * np.cos(th), -s * np.sin(th), rng.uniform(-600, 600), s * np.sin(th), s * np.cos(th),
             rng.uniform(-80, 80), rng.uniform(-3e-4, 3e-4), rng.uniform(-3e-4, 3e-4), rng.uniform(0.5, 2)]
        out.append(((int(v) for v in rng.integers(1, 900, 4)), h))
    # Pure translations by integers (and within rounding of integers): the
    # corners snap, so the canvas is exactly the union box.
    for dx, dy in ((37, -12), (-5, 9), (0, 0)):
        out.append(((64, 48, 64, 48), [1, 0, dx, 0, 1, dy, 0, 0, 1]))
        out.append(((64, 48, 64, 48), [1, 0, dx + 3e-11, 0, 1, dy - 2e-11, 0, 0, 1]))
    out.append(((64, 48, 64, 48), [2, 0, 0, 0, 2, 0, 0, 0, 2]))       # a scaled identity
    out.append(((64, 48, 40, 30), [1, 0, 5, 0, 1, 5, 0.05, 0, 1]))    # horizon through B: mixed signs
    out.append(((64, 48, 40, 30), [1, 0, 5, 0, 1, 5, -1 / 39, 0, 1]))  # a corner on the horizon
    out.append(((64, 48, 64, 48), [1, 2, 3, 2, 4, 6, 1, 2, 3]))       # singular
    out.append(((64, 48, 64, 48), [1, 0, 0, 0, 1e-14, 0, 0, 0, 1]))   # ill-conditioned
    out.append(((64, 48, 64, 48), [0] * 9))                           # zero
    out.append(((64, 48, 64, 48), [1, 0, float("nan"), 0, 1, 0, 0, 0, 1]))
    out.append(((4000, 4000, 4000, 4000), [1, 0, 3900, 0, 1, 100, 0, 0, 1]))  # 7900 x 4100 > budget
    out.append(((1, 1, 1, 1), [1, 0, 0, 0, 1, 0, 0, 0, 1]))           # single pixels
    out.append(((0, 48, 64, 48), [1, 0, 0, 0, 1, 0, 0, 0, 1]))        # invalid size
    return [(tuple(d), h) for d, h in out]


def test_product_geometry_matches_reference_build():
    R = ref_lib()
    gi = (C.c_int * 8)()
    inv = (C.c_double * 9)()
    seen = set()
    for (wa, ha, wb, hb), h in cases():
        st, g = product(wa, ha, wb, hb, h)
        rs = R.ref_s6_prepare_geometry(wa, ha, wb, hb, dbl(h), gi, inv)
        seen.add(st)
        assert st == rs, ((wa, ha, wb, hb), h, st, rs)
        if st == 0:
            assert [g.width, g.height, g.origin_x, g.origin_y, g.width_a, g.height_a, g.width_b, g.height_b] == \
                list(gi), ((wa, ha, wb, hb), h)
            got = np.array(g.inverse[:]).view(np.uint64)
            want = np.array(inv[:]).view(np.uint64)
            assert np.array_equal(got, want), ((wa, ha, wb, hb), h)  # bit patterns, signed zeros included
    assert {0, 1, 2, 5} <= seen  # OK, INVALID_ARGUMENT, INSUFFICIENT_WORKSPACE and DEGENERATE all occur


def test_product_geometry_matches_oracle_and_budget():
    for (wa, ha, wb, hb), h in cases():
        st, g = product(wa, ha, wb, hb, h)
        ost, og = O.prepare_geometry(wa, ha, wb, hb, h)
        assert st == ost
        if st == 0:
            assert (g.width, g.height, g.origin_x, g.origin_y) == (og.width, og.height, og.origin_x, og.origin_y)
            assert list(g.inverse) == list(og.inverse)
    # A larger budget admits the 7900 x 4100 canvas; the box is the union.
    st, g = product(4000, 4000, 4000, 4000, [1, 0, 3900, 0, 1, 100, 0, 0, 1], max_pixels=1 << 26)
    assert st == 0 and (g.width, g.height, g.origin_x, g.origin_y) == (7900, 4100, 0, 0)
    # Integer translation: the canvas is the union box and the inverse the
    # negated shift (normalised by its largest coefficient, 37).
    st, g = product(64, 48, 64, 48, [1, 0, 37, 0, 1, -12, 0, 0, 1])
    assert st == 0 and (g.width, g.height, g.origin_x, g.origin_y) == (101, 60, 0, -12)
    assert np.allclose(np.array(g.inverse[:]) * 37, [1, 0, -37, 0, 1, 12, 0, 0, 1], atol=1e-12)
