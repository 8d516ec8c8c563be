"""The multiply-high modulo the device hash uses (pbf_types.h, make_fast_mod / fast_mod) agrees
with % for every 32-bit input class: the method is checked here over many divisors and inputs;
the compiled function is checked again at solver creation and by the bit-exact hash parity."""
import numpy as np


def make(d):
    if d <= 1:
        return 0, 0
    l = (d - 1).bit_length()
    return ((((1 << l) - d) << 32) // d + 1), l - 1


def fast_mod(x, d, magic, shift):
    if d <= 1:
        return np.zeros_like(x)
    t = (magic * x) >> np.uint64(32)
    q = (t + ((x - t) >> np.uint64(1))) >> np.uint64(shift)
    return x - q * np.uint64(d)


def test_fast_mod_matches_modulo():
    rng = np.random.default_rng(0)
    divisors = [1, 2, 3, 7, 97, 128, 3_456_000, 128_000_000, 1_024_000_000, 2**31 - 1, 2**31,
                2**31 + 1, 2**32 - 1] + [int(v) for v in rng.integers(2, 2**32 - 1, 200)]
    xs = np.concatenate([np.array([0, 1, 2**31 - 1, 2**31, 2**32 - 2, 2**32 - 1], np.uint64),
                         rng.integers(0, 2**32, 20000, dtype=np.uint64)])
    for d in divisors:
        magic, shift = make(d)
        assert magic < 2**32
        x = np.concatenate([xs, np.array([d - 1, d, d + 1, 2 * d - 1, 2 * d], np.uint64) % np.uint64(2**32)])
        got = fast_mod(x, d, np.uint64(magic), shift)
        np.testing.assert_array_equal(got, x % np.uint64(d), err_msg=f"d={d}")
