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
