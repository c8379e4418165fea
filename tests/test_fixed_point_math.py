"""The fixed-point conversion the LDS current block uses in fixed-point mode:
for |v * s| < 2^51 (s a power of two), the double t = v * s + 1.5 * 2^52,
rounded once to the nearest double, holds round-half-even(v * s) in its low
mantissa bits, so bits(t) - bits(1.5 * 2^52) is that integer in two's
complement.  Checked in exact rational arithmetic (a single rounding, as a
fused multiply-add gives) against Python's exact rounding, across the range
the kernel allows and at the ties."""
import struct
from fractions import Fraction

import numpy as np

MAGIC = 1.5 * 2.0 ** 52


def bits(x):
    return struct.unpack("<q", struct.pack("<d", x))[0]


def fixed_point(v, s):
    t = float(Fraction(v) * Fraction(s) + Fraction(MAGIC))  # one correct rounding
    return bits(t) - bits(MAGIC)


def test_magic_number_conversion_is_exact_rounding():
    rng = np.random.default_rng(0)
    for e in (-40, -3, 0, 7):
        s = 2.0 ** (50 - e)  # the kernel's scale for a bound of 2^e
        v = rng.uniform(-1, 1, 2000) * 2.0 ** e
        for x in v:
            want = round(Fraction(x) * Fraction(s))  # round half to even
            assert fixed_point(float(x), s) == want


def test_ties_and_extremes():
    s = 1.0
    for k in (0, 1, -1, 2 ** 50 - 1, -(2 ** 50 - 1), 12345):
        for frac in (Fraction(1, 2), Fraction(-1, 2), Fraction(1, 4), Fraction(3, 4)):
            x = Fraction(k) + frac
            if abs(x) >= 2 ** 50:
                continue
            xf = float(x)
            assert Fraction(xf) == x
            assert fixed_point(xf, s) == round(x)
    # sums of converted contributions are exact integers: order does not matter
    rng = np.random.default_rng(1)
    v = rng.standard_normal(5000)
    s = 2.0 ** 47
    q = [fixed_point(float(x), s) for x in v]
    assert sum(q) == sum(reversed(q))
    assert abs(sum(q) / s - float(np.sum(v))) < 5000 * 0.5 / s + 1e-12
