"""The split-K reduction of the split 16-bit multiply (mixed.hip,
split_reduce_kernel) stores each split's raw FP32 sum and applies
alpha / (scaleA scaleB) to the FP64 sum of the splits; storing the scaled
FP64 values and summing those instead gives the same result bit for bit,
because the operand scales are powers of two (2^(15 - e) from the half
split's range rule), so scaling is exact and commutes with the rounding of
each FP64 addition.  Host check of that arithmetic on random FP32 sums over
the exponent range of the iteration's operands (split order kept, as in the
kernel), the fixed-order reduction of both forms compared element by element.
"""

import numpy as np
import pytest


def _reduce_scaled(parts, unscale, alpha):
    # the former form: each partial scaled in FP64, then summed
    acc = np.zeros(parts.shape[1:], np.float64)
    for p in parts:
        acc = acc + unscale * p.astype(np.float64)
    return alpha * acc


def _reduce_raw(parts, unscale, alpha):
    # split_reduce_kernel: raw FP32 sums added in FP64, one factor at the end
    acc = np.zeros(parts.shape[1:], np.float64)
    for p in parts:
        acc = acc + p.astype(np.float64)
    return (alpha * unscale) * acc


@pytest.mark.parametrize("splits", [1, 2, 8, 16])
@pytest.mark.parametrize("ea,eb", [(15, 15), (40, -3), (-20, 30)])
def test_raw_fp32_partials_reduce_to_the_same_bits(splits, ea, eb):
    rng = np.random.default_rng(splits * 101 + ea)
    # FP32 sums of products of half-split operands scaled into [2^14, 2^15)
    parts = (rng.standard_normal((splits, 64, 48)) * 2.0 ** 28).astype(np.float32)
    parts *= np.float32(2.0) ** rng.integers(-6, 7, size=parts.shape).astype(np.float32)
    unscale = 1.0 / (2.0 ** ea * 2.0 ** eb)
    for alpha in (1.0, -0.5, 0.37, -1.0):
        a = _reduce_scaled(parts, unscale, alpha)
        b = _reduce_raw(parts, unscale, alpha)
        assert np.array_equal(a, b)


def test_a_scale_that_is_not_a_power_of_two_would_differ():
    rng = np.random.default_rng(3)
    parts = (rng.standard_normal((8, 64, 48)) * 2.0 ** 28).astype(np.float32)
    a = _reduce_scaled(parts, 1.0 / 3.0, 0.37)
    b = _reduce_raw(parts, 1.0 / 3.0, 0.37)
    assert not np.array_equal(a, b)
