"""The int64 fixed-point rounding of the slot-order kernels' receiver sums
(smh_batch_slots.h: fixed_scale64 / fixed_bits64), restated on the host:
adding 1.5 * 2^52 to a double u with |u| <= 2^51 rounds u to the nearest
integer and leaves kMagic64Bits + rint(u) in its bit pattern, so summing the
patterns as uint64 (wrapping) and subtracting kMagic64Bits once per term
gives the exact integer sum."""
import numpy as np

K_MAGIC = 6755399441055744.0
K_MAGIC_BITS = np.uint64(0x4338000000000000)


def _bits(u):
    return (np.asarray(u, dtype=np.float64) + K_MAGIC).view(np.uint64)


def test_magic_rounding_matches_rint():
    rng = np.random.default_rng(3)
    for scale in (1.0, 1e3, 2.0 ** 30, 2.0 ** 51):
        u = rng.uniform(-scale, scale, 20000)
        got = (_bits(u) - K_MAGIC_BITS).view(np.int64)
        assert np.array_equal(got, np.rint(u).astype(np.int64)), scale
    # round half to even, as rint
    u = np.array([0.5, 1.5, 2.5, -0.5, -1.5, 2.0 ** 51 - 0.5])
    assert np.array_equal((_bits(u) - K_MAGIC_BITS).view(np.int64), np.rint(u).astype(np.int64))


def test_wrapped_sum_minus_bias_is_exact():
    rng = np.random.default_rng(4)
    e_n = 14  # 2^14 terms per word, each within 2^(62 - 14)
    n = 2 ** e_n
    u = rng.uniform(-2.0 ** (62 - e_n), 2.0 ** (62 - e_n), n)
    u = np.clip(u, -2.0 ** 51, 2.0 ** 51)
    with np.errstate(over='ignore'):
        acc = np.add.reduce(_bits(u), dtype=np.uint64)  # wraps modulo 2^64
        total = (acc - np.uint64(n) * K_MAGIC_BITS).view(np.int64)
    want = int(sum(int(v) for v in np.rint(u).astype(np.int64)))
    assert int(total) == want


def test_scale_keeps_every_term_in_range():
    # sh = 2^k with k = min(51, 62 - eN) - (floor(log2 V) + 1): |v| sh <= 2^min(51, 62 - eN) for |v| <= V
    rng = np.random.default_rng(5)
    for _ in range(200):
        V = 10.0 ** rng.uniform(-20, 30)
        e_n = int(rng.integers(0, 20))
        k = min(51, 62 - e_n) - (int(np.floor(np.log2(V))) + 1)
        k = max(-126, min(k, 126))
        assert V * 2.0 ** k <= 2.0 ** min(51, 62 - e_n)
        assert 2.0 ** e_n * V * 2.0 ** k <= 2.0 ** 62
