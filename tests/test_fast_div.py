"""The segment kernels decode a membership index t (t < 2^31) into (det, slot)
with floor(t / d) computed as the high 64 bits of t * ceil(2^64 / d) (d >= 2;
d is a count of removal patterns per det, at most C(64, 2)).  Checked
against integer division over every divisor the layouts can use, at the
dividends nearest to each multiple of d and at random ones up to 2^33."""
import numpy as np


def _div_mul(d):
    return (2**64 - 1) // d + 1


def test_multiply_high_division_is_exact():
    rng = np.random.default_rng(7)
    for d in range(2, 2017):
        mul = _div_mul(d)
        ts = [int(x) for x in rng.integers(0, 2**33, 64)]
        top = (2**31 // d) * d
        ts += [top - 1, top, top + 1, 2**31 - 1, 2**33 - 1, d - 1, d, d + 1, 0]
        for t in ts:
            assert (t * mul) >> 64 == t // d, (t, d)
