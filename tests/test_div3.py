"""The division by 3 of the 3D local solve (csrc/eikonal_device.hpp, div3):
q0 = RN(x y) with y = RN(1/3), r = x - 3 q0, q = RN(q0 + r y), the residual
clamped to at most 2^1023 (IEEE minNum, so NaN becomes 2^1023).

Pinned on the CPU in exact rational arithmetic against IEEE division
(Python's float division is correctly rounded): the residual is exactly
representable (so the FMA computing it is exact) and far below the clamp,
and q equals x / 3 for
normal, subnormal and boundary binary64 inputs, +-inf and NaN; -0 gives +0.
The device implementation is checked against IEEE division on the GPU
(tests/test_gpu_parity.py::test_device_arith_is_ieee).
"""
import math
from fractions import Fraction

import numpy as np

Y = float.fromhex("0x1.5555555555555p-2")
CLAMP = 2.0 ** 1023
BIG = float.fromhex("0x1.fffffffffffffp+1023")


def rn(q):
    """Fraction -> nearest binary64 (ties to even), as an FMA rounds."""
    return float(q)


def div3(x):
    q0 = x * Y
    if math.isinf(x) or math.isnan(x):
        r = math.nan  # inf - inf
    else:
        r_exact = Fraction(x) - 3 * Fraction(q0)
        r = float(r_exact)
        assert Fraction(r) == r_exact, x  # the FMA's residual is exact
        assert abs(r) <= 2.0 ** 972, x  # the clamp never changes a finite residual
    r = CLAMP if math.isnan(r) else min(r, CLAMP)  # IEEE minNum: the number survives
    if math.isinf(q0) or math.isnan(q0):
        return q0 + r * Y  # inf + finite, or NaN
    return rn(Fraction(q0) + Fraction(r) * Fraction(Y))


def _inputs():
    rng = np.random.default_rng(7)
    bits = rng.integers(0, 2 ** 63, 20000, dtype=np.int64)
    xs = list(bits.view(np.float64))
    sub = rng.integers(1, 2 ** 52, 10000, dtype=np.int64)  # subnormals
    xs += list(sub.view(np.float64))
    low = rng.integers(1, 60, 10000) << 52 | rng.integers(0, 2 ** 52, 10000)  # lowest binades
    xs += list(np.asarray(low, dtype=np.int64).view(np.float64))
    xs += [5e-324, 1e-310, 2.0 ** -1022, 3 * 2.0 ** -1022, 2.0 ** -1074 * 3, 1.0, 3.0,
           BIG, -BIG, 1e308, 0.1, 2.0 / 3.0]
    xs = [x for x in xs if math.isfinite(x)]
    return xs + [-x for x in xs[:5000]]


def test_three_operation_division_is_correctly_rounded():
    for x in _inputs():
        q = div3(float(x))
        assert q == float(x) / 3.0 and math.copysign(1, q) == math.copysign(1, float(x) / 3.0), x


def test_special_values():
    assert div3(math.inf) == math.inf and div3(-math.inf) == -math.inf
    assert math.isnan(div3(math.nan))
    assert div3(0.0) == 0.0 and math.copysign(1, div3(0.0)) == 1
    assert div3(-0.0) == 0.0 and math.copysign(1, div3(-0.0)) == 1  # the one exception
