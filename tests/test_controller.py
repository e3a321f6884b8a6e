"""Controller (legacy-v1 rule policy) and plan_retry, checked against the
reference's own controller compiled from its sources (oracle/_ref) when that
build is present, and against the ladder's closed-form rules otherwise."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
REF = ROOT / "oracle" / "_ref" / "libref.so"


def hres(status, matches, inliers, coverage, residual):
    from cs_stitch import _ffi as F
    h = F.HomographyResult()
    for i in range(9):
        h.H[i] = h.H_s[i] = 1.0 if i % 4 == 0 else 0.0
    h.status = status
    h.num_matches = matches
    h.num_inliers = inliers
    h.inlier_ratio = inliers / matches if matches else 0.0
    h.inlier_coverage = coverage
    h.reproj_residual = residual
    return h


def scenarios():
    rng = np.random.default_rng(4)
    out = []
    for _ in range(300):
        ga, gb = rng.uniform(0, 0.15, 2)
        n = int(rng.integers(1, 6))
        fbs, st = [], []
        for _ in range(n):
            kind = rng.integers(0, 10)
            if kind == 0:
                m = int(rng.integers(0, 4))
                fbs.append(hres(4, m, 0, 0, float("nan")))
            elif kind == 1:
                fbs.append(hres(5, int(rng.integers(4, 50)), 0, 0, float("nan")))
            else:
                m = int(rng.integers(4, 400))
                k = int(rng.integers(4, m + 1))
                fbs.append(hres(0, m, k, float(rng.uniform(0, 0.5)), float(rng.uniform(0, 5))))
            st.append(0)
        out.append((ga, gb, fbs, st))
    return out


def test_plan_retry_rules():
    from cs_stitch import _ffi as F
    L = F.load()
    plan = (C.c_int * 4)()
    assert L.cs_plan_retry(.25, 500, 128, 0, .5, 1000, 128, 0, plan) == 0 and list(plan) == [1, 0, 0, 1]
    assert L.cs_plan_retry(.5, 1000, 128, 0, .5, 2000, 128, 0, plan) == 0 and list(plan) == [2, 0, 0, 1]
    assert L.cs_plan_retry(.5, 1000, 128, 0, .5, 1000, 256, 0, plan) == 0 and list(plan) == [5, 1, 128, 0]
    assert L.cs_plan_retry(.5, 1000, 256, 0, .5, 1000, 128, 0, plan) == 0 and list(plan) == [5, 0, 0, 0]
    assert L.cs_plan_retry(.5, 1000, 256, 0, .5, 1000, 256, 0, plan) == 1
    assert L.cs_plan_retry(.5, 1000, 256, 0, .5, 1000, 512, 1, plan) == 1


@pytest.mark.skipif(not REF.exists(), reason="reference controller build (oracle/_ref) not present")
def test_legacy_walk_matches_reference_controller():
    from cs_stitch import pipeline as P
    ref = C.CDLL(str(REF))
    ref.ref_legacy_walk.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    ref.ref_plan_retry.argtypes = [C.c_double, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                   C.c_void_p]
    for blind in (0, 1):
        ctl = P.Controller(blind=bool(blind))
        for ga, gb, fbs, st in scenarios():
            mine = ctl.walk(ga, gb, 640, 480, 600, 400, fbs, st)
            n = len(fbs)
            arr = (P.F.HomographyResult * n)(*fbs)
            sts = (C.c_int * n)(*st)
            trace = (C.c_double * (5 * (n + 2)))()
            rows = ref.ref_legacy_walk(blind, ga, gb, 640, 480, 600, 400, arr, sts, n, trace)
            assert mine == [tuple(trace[5 * i:5 * i + 5]) for i in range(rows)]
    from cs_stitch import _ffi as F
    L = F.load()
    a, b = (C.c_int * 4)(), (C.c_int * 4)()
    ladder = [(.25, 500, 128), (.5, 1000, 128), (.5, 1000, 256), (.75, 2000, 512), (1., 2000, 1024), (1., 2000, 512)]
    for x in ladder:
        for y in ladder:
            r1 = L.cs_plan_retry(*x, 0, *y, 0, a)
            r2 = ref.ref_plan_retry(*x, 0, *y, 0, b)
            assert r1 == r2 and (r1 != 0 or list(a) == list(b))
