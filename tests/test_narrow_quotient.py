"""The narrow phase splits a flat test index k into (own, other) as
floor(k / n) = trunc(((float)k + 0.5f) * (1.0f / n)) in FP32 for k < 2^20
(rk_nl.hip, quot_small).  The identity holds because k + 0.5 is exact in
FP32, the product's relative error stays below the 0.5 / n distance from
(k + 0.5) / n to an integer, and the kernel is built without fast-math (so
1.0f / n and the product are correctly rounded, as numpy's float32 ops are).
Checked here exhaustively over k < 2^20 and n <= 2048.  n is a count of one
shape's vertices, edges or faces, and the narrow phase's 64 KiB LDS budget
(12 bytes per 64-test word of near bits, n^2 tests for vv or ee alone) keeps
every such count below 600."""
import numpy as np

K_LIMIT = 1 << 20


def test_fp32_quotient_matches_integer_division():
    k = np.arange(K_LIMIT, dtype=np.int64)
    kf = k.astype(np.float32) + np.float32(0.5)
    for n in range(1, 2049):
        inv = np.float32(1.0) / np.float32(n)
        q = (kf * inv).astype(np.int64)  # truncation toward zero, k >= 0
        bad = np.nonzero(q != k // n)[0]
        assert bad.size == 0, (n, int(bad[0]))


def test_planted_mistake_is_caught():
    """Without the half offset the quotient is wrong at exact multiples of
    some n (the product rounds below the integer)."""
    k = np.arange(K_LIMIT, dtype=np.int64)
    kf = k.astype(np.float32)
    wrong = 0
    for n in range(1, 200):
        inv = np.float32(1.0) / np.float32(n)
        wrong += int(np.count_nonzero((kf * inv).astype(np.int64) != k // n))
    assert wrong > 0
