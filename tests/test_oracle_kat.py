"""Known-answer checks of the RNG boundary the reference depends on.

glibc rand() and libstdc++ default_random_engine / normal_distribution are not
vendored in the reference; these values (measured on glibc 2.35 / GCC 11.4)
pin them, both as the oracle calls them and as the product's host layer draws
them (glibc random_r with a private state).
"""
import ctypes as C

import numpy as np
import pytest


def test_glibc_rand_seed0(oracle):
    out = np.empty(6, np.int32)
    oracle.lib().oracle_glibc_rand(0, 6, out.ctypes.data_as(C.POINTER(C.c_int32)))
    assert out.tolist() == [1804289383, 846930886, 1681692777, 1714636915, 1957747793, 424238335]


def test_glibc_rand_seed1_equals_seed0(oracle):
    a = np.empty(50, np.int32)
    b = np.empty(50, np.int32)
    oracle.lib().oracle_glibc_rand(0, 50, a.ctypes.data_as(C.POINTER(C.c_int32)))
    oracle.lib().oracle_glibc_rand(1, 50, b.ctypes.data_as(C.POINTER(C.c_int32)))
    assert np.array_equal(a, b)


def test_glibc_rand_rank1_seed(oracle):
    out = np.empty(3, np.int32)
    oracle.lib().oracle_glibc_rand(10000, 3, out.ctypes.data_as(C.POINTER(C.c_int32)))
    assert out.tolist() == [595422259, 506549998, 979315720]


def test_glibc_type3_recurrence(oracle):
    """The published TYPE_3 additive generator: r[i] = r[i-3] + r[i-31] (mod 2^32),
    output r[i] >> 1, state from the 16807 LCG with 310 discarded outputs."""
    seed = 12345
    r = [0] * 344
    r[0] = seed
    for i in range(1, 31):
        hi, lo = divmod(r[i - 1], 127773)
        w = 16807 * lo - 2836 * hi
        if w < 0:
            w += 2147483647
        r[i] = w
    for i in range(31, 34):
        r[i] = r[i - 31]
    for i in range(34, 344):
        r[i] = (r[i - 31] + r[i - 3]) & 0xFFFFFFFF
    want = []
    for i in range(344, 344 + 100):
        r.append((r[i - 31] + r[i - 3]) & 0xFFFFFFFF)
        want.append(r[i] >> 1)
    out = np.empty(100, np.int32)
    oracle.lib().oracle_glibc_rand(seed, 100, out.ctypes.data_as(C.POINTER(C.c_int32)))
    assert out.tolist() == want


def test_minstd_rand0(oracle):
    out = np.empty(10, np.int64)
    oracle.lib().oracle_minstd(10, out.ctypes.data_as(C.POINTER(C.c_int64)))
    x = 1
    want = []
    for _ in range(10):
        x = (16807 * x) % 2147483647
        want.append(x)
    assert out.tolist() == want
    assert out[0] == 16807 and out[1] == 282475249


def test_normal_distribution_kat(oracle):
    out = np.empty(4, np.float64)
    oracle.lib().oracle_std_normal(3.0, 1.0, 4, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert out.tolist() == [2.8780342158584031, 1.9131819557386427, 3.6842899437965548,
                            1.924810850481971]
    oracle.lib().oracle_std_normal(5.0, 1.0, 2, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert out[:2].tolist() == [4.8780342158584027, 3.9131819557386427]


def test_normal_distribution_polar_method(oracle):
    """libstdc++'s normal_distribution is Marsaglia's polar method over
    generate_canonical<double,53>(minstd_rand0) (two engine calls per uniform);
    pairs are returned second value first."""
    m = 2147483647
    x = 1

    def eng():
        nonlocal x
        x = (16807 * x) % m
        return x

    R = 2147483646.0

    def canon():
        s = float(eng() - 1) + float(eng() - 1) * R
        # 2^62-ish divisor: R*R rounded to double (computed in long double)
        t = float(np.longdouble(R) * np.longdouble(R))
        v = s / t
        return v if v < 1.0 else np.nextafter(1.0, 0.0)

    want = []
    while len(want) < 40:
        while True:
            a = 2.0 * canon() - 1.0
            b = 2.0 * canon() - 1.0
            r2 = a * a + b * b
            if r2 <= 1.0 and r2 != 0.0:
                break
        mult = np.sqrt(-2 * np.log(r2) / r2)
        want += [b * mult, a * mult]
    out = np.empty(40, np.float64)
    oracle.lib().oracle_std_normal(0.0, 1.0, 40, out.ctypes.data_as(C.POINTER(C.c_double)))
    np.testing.assert_allclose(out, want, rtol=0, atol=1e-15)


def test_sqrtf_constants(oracle):
    L = oracle.lib()
    assert float(L.oracle_sqrtf(3.0)) == 1.7320507764816284
    assert float(L.oracle_sqrtf(2.0)) == 1.4142135381698608
    assert float(np.sqrt(np.float32(3.0))) == 1.7320507764816284
