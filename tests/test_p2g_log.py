"""Accuracy of the logarithm used for geometric P2G averaging (p2g_log in
mdoodz7.0_amd/csrc/p2g_common.h), restated operation by operation in IEEE
double arithmetic (fma evaluated exactly with rationals, then rounded once,
as the hardware instruction does) and compared with numpy's log.  The GPU
P2G tests check the device code itself end to end against the oracle."""
import math
import struct
from fractions import Fraction

import numpy as np


def fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def p2g_log(v):
    if not (v >= 2.2250738585072014e-308) or not (v <= 1.7976931348623157e308):
        return math.log(v) if v > 0 else float("nan")
    m, e = math.frexp(v)
    if m < 0.70710678118654752440:
        m = 2.0 * m
        e -= 1
    f = m - 1.0
    s = f / (2.0 + f)
    z = s * s
    p = 1.0 / 21.0
    for c in (19.0, 17.0, 15.0, 13.0, 11.0, 9.0, 7.0, 5.0, 3.0):
        p = fma(p, z, 1.0 / c)
    s2 = s + s
    k = float(e)
    r = fma(s2 * z, p, k * 1.9082149292705877e-10)
    return k * 0.6931471803691238 + (s2 + r)


def ulps(a, b):
    return abs(a - b) / math.ulp(b)


def test_ln2_split():
    # hi has its low 21 mantissa bits clear and hi + lo rounds to ln 2
    hi, lo = 0.6931471803691238, 1.9082149292705877e-10
    assert struct.unpack("<Q", struct.pack("<d", hi))[0] & ((1 << 21) - 1) == 0
    assert hi + lo == math.log(2.0)


def test_p2g_log_within_three_ulp():
    rng = np.random.default_rng(11)
    vals = list(10.0 ** rng.uniform(-300, 300, 1500))          # the full normal range
    vals += list(1.0 + rng.uniform(-0.3, 0.3, 800))           # near 1 (k = 0, small s)
    vals += list(10.0 ** rng.uniform(-3, 3, 800))             # viscosity-like values
    vals += [0.70710678118654752440, 0.7071067811865475, 0.7071067811865476, 2.0, 0.5,
             math.sqrt(2.0), 1e-3, 1e3, 2.2250738585072014e-308, 1.7976931348623157e308]
    worst = 0.0
    for v in vals:
        ref = float(np.log(np.float64(v)))
        worst = max(worst, ulps(p2g_log(v), ref))
    assert p2g_log(1.0) == 0.0
    print("worst error %.2f ulp" % worst)
    assert worst <= 3.0   # numpy's own log is within 1 ulp of the exact value


def test_p2g_log_special_values():
    assert math.isnan(p2g_log(-1.0))
    assert p2g_log(float("inf")) == float("inf")
    assert p2g_log(5e-324) == math.log(5e-324)
