"""The transform passes divide stage indices by the stage sizes with one multiply-high:
q = floor(x m / 2^32) with m = floor(2^32 / d) + 1 (the plan's multipliers, d = 1 handled
apart).  The passes' dividends stay below 2^16 (a tile holds at most 16 lines of 4096) and the
divisors are stage sizes of lengths up to 4096; this checks the identity q = x // d for every
such divisor and dividend."""
import numpy as np


def test_multiply_high_division_is_exact_for_every_stage_size():
    x = np.arange(1 << 16, dtype=np.uint64)
    for d in range(2, 4097):
        m = np.uint64((1 << 32) // d + 1)
        q = (x * m) >> np.uint64(32)
        assert np.array_equal(q, x // np.uint64(d)), d


def test_multiply_high_division_fails_past_its_range():
    """the range matters: a dividend of 2^20 is already wrong for some divisor, so the bound on
    the tile size is what makes the passes exact"""
    x = np.arange(1 << 20, 1 << 21, dtype=np.uint64)
    wrong = False
    for d in (3, 7, 1000, 4095):
        m = np.uint64((1 << 32) // d + 1)
        wrong |= not np.array_equal((x * m) >> np.uint64(32), x // np.uint64(d))
    assert wrong
