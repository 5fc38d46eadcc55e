"""The node terms' table-driven log / exp: the tables the host builds in
80-bit arithmetic, checked against mpmath at 60 digits (no device needed).
Each high + low pair must carry its value to well below one FP64 ulp (the
80-bit arithmetic gives about 2^-64 relative), the Cody-Waite high parts must
have their low 20 significand bits clear (exact products with exponents up
to 2^20), and each c_j must be within one ulp of 1 / (1 + (j + 1/2)/128)."""
import mpmath as mp
import numpy as np

import gpu_determ_amd as G

mp.mp.dps = 60


def low_bits_clear(x, bits):
    u = np.array([x], np.float64).view(np.uint64)[0]
    return int(u) & ((1 << bits) - 1) == 0


def test_log_table():
    T = G.kmath_tables()
    assert T.shape == (520,)
    for j in range(128):
        c = T[j]
        exact_c = 1 / (1 + (mp.mpf(j) + mp.mpf(1) / 2) / 128)
        assert abs(mp.mpf(c) - exact_c) <= mp.mpf(np.spacing(c))
        v = -mp.log(mp.mpf(c))
        assert abs(mp.mpf(T[128 + j]) + mp.mpf(T[256 + j]) - v) <= mp.mpf(2) ** -62 * max(abs(v), mp.mpf(2) ** -8)


def test_exp_table():
    T = G.kmath_tables()
    for j in range(64):
        v = mp.mpf(2) ** (mp.mpf(j) / 64)
        assert abs(mp.mpf(T[384 + j]) + mp.mpf(T[448 + j]) - v) <= mp.mpf(2) ** -62 * v


def test_constants():
    T = G.kmath_tables()
    ln2 = mp.log(2)
    assert low_bits_clear(T[512], 20) and low_bits_clear(T[515], 20)
    assert abs(mp.mpf(T[512]) + mp.mpf(T[513]) - ln2) <= mp.mpf(2) ** -62
    assert abs(mp.mpf(T[514]) - 64 / ln2) <= mp.mpf(np.spacing(T[514]))
    assert abs(mp.mpf(T[515]) + mp.mpf(T[516]) - ln2 / 64) <= mp.mpf(2) ** -68
    assert np.all(T[517:] == 0.0)
