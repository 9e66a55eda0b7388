"""The fast reciprocal of the path kernels (hardware reciprocal plus one
Newton-Raphson step on fused multiply-adds) against IEEE division, over
every float32 bit pattern of the range where the kernels use it: bit for
bit.  Just outside that range (quotients that are denormal) the two differ,
which is why the kernels fall back to the division there."""
import pytest

pytestmark = pytest.mark.gpu


def test_fast_reciprocal_equals_ieee_division_on_normal_range(ctx):
    # 2^-126 <= x < 2^126, positive and negative: 2 x 2,113,929,216 patterns
    for lo, hi in ((0x00800000, 0x7E800000), (0x80800000, 0xFE800000)):
        n, first = ctx.selftest_rcp(lo, hi)
        print(f"rcp selftest [{lo:08x}, {hi:08x}): {n} mismatches {[hex(f) for f in first]}", flush=True)
        assert n == 0


def test_fast_reciprocal_differs_outside_it(ctx):
    # x >= 2^126 (quotients below the normal range) and denormal x
    n_big, first = ctx.selftest_rcp(0x7E800000, 0x7F000000)
    assert n_big > 0 and all(0x7E800000 <= f < 0x7F000000 for f in first)
    n_small, _ = ctx.selftest_rcp(0x00000001, 0x00800000)
    assert n_small > 0
    assert ctx.selftest_rcp(5, 5) == (0, [])
