"""The post passes divide small integers (an 8-bit channel by 255, a BMFR block
coordinate by the block size - 1, the coordinate features of bmfrPre.comp:36-43 that
the post pass recomputes, and the UNORM8 reads of the albedo and sample-count images) by a reciprocal multiply and
one FMA correction, `div_small<kD>` in csrc/image_ops.hpp.  This checks, in exact
rational arithmetic, that the sequence q = RN(c * RN(1/d)), e = RN(c - q*d),
result = RN(e * RN(1/d) + q) in binary32 equals the IEEE quotient RN(c / d) for every
c in [0, 255] and every d in [1, 255], the block sizes' divisors among them."""
from fractions import Fraction

import numpy as np


def round_f32(x: Fraction) -> np.float32:
    """x rounded to the nearest binary32 value, ties to even."""
    f = np.float32(float(x))
    cands = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
    best = min(cands, key=lambda c: (abs(Fraction(float(c)) - x), int(np.float32(c).view(np.uint32)) & 1))
    return np.float32(best)


def fma_f32(a, b, c) -> np.float32:
    return round_f32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def div_small(c: int, d: int) -> np.float32:
    r = np.float32(1.0) / np.float32(d)   # the compile-time constant 1/d, rounded to nearest
    cf = np.float32(c)
    q = cf * r                            # binary32 product, rounded to nearest
    e = fma_f32(-q, np.float32(d), cf)
    return fma_f32(e, r, q)


def test_round_f32_is_correct_rounding():
    # halfway between 1 and the next binary32 value rounds to the even one (1.0);
    # just above it rounds up
    ulp = Fraction(1, 2 ** 23)
    assert round_f32(1 + ulp / 2) == np.float32(1.0)
    assert round_f32(1 + ulp / 2 + Fraction(1, 2 ** 60)) == np.nextafter(np.float32(1), np.float32(2))
    assert round_f32(1 + 3 * ulp / 2) == np.float32(1 + 2 * 2.0 ** -23)


def test_div_small_equals_ieee_division():
    bad = []
    for d in range(1, 256):
        for c in range(256):
            if div_small(c, d) != np.float32(c) / np.float32(d):
                bad.append((c, d))
    assert bad == []


def test_plain_reciprocal_product_is_not_enough():
    # without the correction step c * RN(1/d) misses the IEEE quotient for some
    # 8-bit values (so the correction is doing work)
    r = np.float32(1.0) / np.float32(255)
    misses = [c for c in range(256) if np.float32(c) * r != np.float32(c) / np.float32(255)]
    assert misses
