"""Exact-arithmetic checks of closed forms the S6 kernels use in place of
tables or divisions (crossstage-stitch_amd/csrc/s6.hip).

unit_byte: A's level-0 colour is float(v / 255.0) for a byte v
(src/s6/s6.cu:171, the FP64 quotient rounded to FP32). The kernels form it
as q = v * RN(1/255); q + RN(v - 255 q) * RN(1/255), both steps fused
multiply-adds in FP32. Every intermediate is evaluated here in exact
rational arithmetic and rounded to FP32 once, as a fused multiply-add does,
for all 256 bytes."""
from fractions import Fraction

import numpy as np


def rn32(x: Fraction) -> np.float32:
    """Round an exact rational to the nearest FP32 (ties to even)."""
    f = np.float32(float(x))  # within an FP32 ulp of x (possibly on the wrong side)
    best = None
    for c in (np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))):
        d = abs(Fraction(float(c)) - x)
        if best is None or d < best[0] or (d == best[0] and int(c.view(np.uint32)) % 2 == 0):
            best = (d, c)
    return best[1]


def fma32(a, b, c) -> np.float32:
    return rn32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def mul32(a, b) -> np.float32:
    return rn32(Fraction(float(a)) * Fraction(float(b)))


def test_rn32_rounds_ties_to_even_and_halfway_cases():
    one = np.float32(1.0)
    ulp = Fraction(1, 2 ** 23)
    assert rn32(Fraction(1) + ulp / 2) == one                      # tie -> even (1.0)
    assert rn32(Fraction(1) + 3 * ulp / 2) == np.float32(1 + 2 * 2.0 ** -23)  # tie -> even
    assert rn32(Fraction(1) + ulp / 2 + Fraction(1, 2 ** 60)) == np.nextafter(one, np.float32(2))


def test_unit_byte_equals_the_fp64_quotient_rounded_to_fp32_for_every_byte():
    inv = np.float32(1.0) / np.float32(255.0)  # RN32(1/255)
    assert Fraction(float(inv)) == Fraction(float(rn32(Fraction(1, 255))))
    plain_product_misses = 0
    for v in range(256):
        want = np.float32(v / 255.0)  # the FP64 quotient, then rounded to FP32
        assert Fraction(float(want)) == Fraction(float(rn32(Fraction(v, 255)))), v  # no double-rounding case
        fv = np.float32(v)
        q = mul32(fv, inv)
        r = fma32(-q, np.float32(255.0), fv)
        got = fma32(r, inv, q)
        assert got == want, (v, got, want)
        plain_product_misses += q != want
    # The correction step matters: the product alone misses about half.
    assert plain_product_misses > 100


# Ownership of overlap pixels (crossstage-stitch_amd/csrc/s6.hip, a_owns):
# A owns a pixel where sa >= sb (src/s6/s6.cu:206-207), each score a minimum
# of correctly rounded quotients d / n (s6.cu:170). The kernel compares
# products d * RN(1 / n) instead, trusting their order when they differ by
# more than 2^-48 of the larger, which rests on each product lying within
# 3 * 2^-53 (relative) of the rounded quotient.
SIZES = [1, 2, 3, 7, 480, 640, 720, 1080, 1280, 1920, 2160, 3840, 4095, 4096, 8191, 16384]


def test_reciprocal_products_within_three_half_ulps_of_the_quotients():
    for n in SIZES:
        d = np.arange(1, n + 1, dtype=np.float64) - 0.5   # every d the kernel forms: 0.5 .. n - 0.5
        exact = d / np.float64(n)                         # IEEE: the correctly rounded quotient
        fast = d * (np.float64(1.0) / np.float64(n))
        assert np.all(np.abs(fast - exact) <= 3 * 2.0 ** -53 * exact), n
    # The bound is tight enough to matter: the products alone miss often.
    d = np.arange(1, 3841, dtype=np.float64) - 0.5
    assert np.count_nonzero(d * (1.0 / 3840.0) != d / 3840.0) > 100


def _scores(xa, wa, sya, xb, yb, wb, hb, fast):
    if fast:
        qa = min(min(xa + 0.5, wa - 0.5 - xa) * (1.0 / wa), sya)
        qb = min(min(xb + 0.5, wb - 0.5 - xb) * (1.0 / wb), min(yb + 0.5, hb - 0.5 - yb) * (1.0 / hb))
    else:
        qa = min(min(xa + 0.5, wa - 0.5 - xa) / wa, sya)
        qb = min(min(xb + 0.5, wb - 0.5 - xb) / wb, min(yb + 0.5, hb - 0.5 - yb) / hb)
    return qa, qb


def test_fast_ownership_agrees_with_the_quotients_or_defers():
    rng = np.random.default_rng(5)
    deferred = 0
    for _ in range(20000):
        wa, wb, hb, ha = (int(rng.choice(SIZES[4:])) for _ in range(4))
        xa = float(rng.integers(0, wa))
        ya = float(rng.integers(0, ha))
        sya = min(ya + 0.5, ha - 0.5 - ya) / ha
        # B positions: general, on the integer grid (translations) and mirrored
        # onto A's score (exact ties).
        kind = rng.integers(3)
        xb = float(rng.uniform(0, wb - 1)) if kind == 0 else float(rng.integers(0, wb))
        yb = float(rng.uniform(0, hb - 1)) if kind == 0 else float(rng.integers(0, hb))
        if kind == 2:
            wb = wa
            xb = wb - 1 - xa
        qa, qb = _scores(xa, wa, sya, xb, yb, wb, hb, True)
        sa, sb = _scores(xa, wa, sya, xb, yb, wb, hb, False)
        if abs(qa - qb) > 2.0 ** -48 * max(qa, qb):
            assert (qa - qb > 0) == (sa >= sb) and sa != sb
        else:
            deferred += 1
    assert deferred < 20000 // 3 + 500  # mostly the planted ties


# B's sample position snapping (crossstage-stitch_amd/csrc/s6.hip): the
# reference's rule (src/s6/s6.cu:148-155) snaps x to the nearest integer when
# within max(tol, 16 eps max(1, |x|)) and then requires 0 <= x <= n - 1. The
# warp's short form uses tol alone when tol >= 16 eps max(1, n); the two must
# give the same validity and snapped value wherever the short form is used.
EPS = 2.2204460492503131e-16


def _snap_full(x, n, tol):
    if not np.isfinite(x):
        return False, x
    nearest = np.rint(x)
    roundoff = 16 * EPS * max(1.0, abs(x))
    if abs(x - nearest) <= max(tol, roundoff):
        x = nearest
    return (x >= 0 and x <= n - 1.0), x


def _snap_tol(x, n, tol):
    with np.errstate(invalid="ignore"):
        nearest = np.rint(x)
        if abs(x - nearest) <= tol:
            x = nearest
        return bool(x >= 0 and x <= n - 1.0), x


def test_short_snap_agrees_with_the_reference_rule():
    rng = np.random.default_rng(11)
    checked = 0
    for n in (1, 2, 480, 1080, 3840, 16384, 28000):
        for tol in (1e-10, 1e-6, 0.25, 0.49, 16 * EPS * max(1.0, n)):
            if not tol >= 16 * EPS * max(1.0, float(n)):
                continue
            xs = [0.0, -0.0, n - 1.0, float(n), -tol, n - 1.0 + tol, -2 * tol, n - 1.0 + 2 * tol,
                  np.nextafter(-tol, -1.0), np.nextafter(n - 1 + tol, np.inf), np.inf, -np.inf, np.nan,
                  1e300, -1e300, n * 1.5, float(n) + 0.49]
            ks = rng.integers(-3, n + 3, 400).astype(np.float64)
            for k in ks:
                for d in (0.0, tol, -tol, np.nextafter(tol, 1.0), -np.nextafter(tol, 1.0), 0.5, 1e-3):
                    xs.append(k + d)
            xs += list(rng.uniform(-n, 2 * n, 400))
            for x in xs:
                vf, xf = _snap_full(float(x), n, tol)
                vt, xt = _snap_tol(float(x), n, tol)
                assert vf == vt, (n, tol, x)
                if vf:
                    assert xf == xt, (n, tol, x)
                checked += 1
    assert checked > 10000
