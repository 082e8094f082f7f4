"""Pins the oracle's primitives: the Philox stream against the published
Random123 known-answer vectors, and its own FP64 log/exp/pow/sin/cos
against the C library."""
import math

import numpy as np

import oracle_bind as ob

L = ob.lib()


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 10 rounds
    assert ob.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert ob.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert ob.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_uniform_mapping_is_exact():
    assert L.or_uniform_from(0, 0) == 0.0
    assert L.or_uniform_from(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 2.0 ** -53
    assert L.or_uniform_from(0x80000000, 0) == 0.5


def test_stream_statistics():
    u = ob.stream(42, ob.OR_RNG_COLLIDE, 7, 3, 0, 200000)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 4 * math.sqrt(1 / 12 / u.size)
    # serial correlation of consecutive draws
    c = np.corrcoef(u[:-1], u[1:])[0, 1]
    assert abs(c) < 4 / math.sqrt(u.size)
    # different streams differ
    v = ob.stream(42, ob.OR_RNG_COLLIDE, 8, 3, 0, 1000)
    assert not np.array_equal(u[:1000], v)
    g = ob.stream_gauss(42, ob.OR_RNG_MOVE, 1, 0, 0, 200000)
    assert abs(g.mean()) < 4 / math.sqrt(g.size)
    assert abs(g.var() - 1.0) < 4 * math.sqrt(2 / g.size)
    assert abs((g ** 4).mean() - 3.0) < 0.05


def _ulp_err(got, want):
    return abs(got - want) / math.ulp(abs(want)) if want != 0 else abs(got)


def test_log_exp_pow_against_libm():
    rng = np.random.default_rng(0)
    xs = np.concatenate([np.exp(rng.uniform(-700, 700, 20000)), rng.uniform(0, 1, 20000), [1e-310, 5e-324, 1.0]])
    worst = max(_ulp_err(L.or_log(x), math.log(x)) for x in xs if x != 1.0)
    assert worst <= 2.0, worst
    assert L.or_log(1.0) == 0.0
    ys = np.concatenate([rng.uniform(-745, 709, 20000), rng.uniform(-1, 1, 20000)])
    worst = max(_ulp_err(L.or_exp(y), math.exp(y)) for y in ys if math.exp(y) > 1e-300)
    assert worst <= 2.0, worst
    for x, y in [(2.0, 0.5), (1e-20, -0.31), (3.7, 2.0), (0.25, 1 / 1.4)]:
        assert abs(L.or_pow(x, y) - x ** y) <= 1e-15 * x ** y
    assert L.or_pow(0.0, 0.5) == 0.0 and L.or_pow(5.0, 0.0) == 1.0


def test_sincos2pi_against_libm():
    rng = np.random.default_rng(1)
    us = np.concatenate([rng.uniform(0, 1, 50000), [0.0, 0.125, 0.25, 0.5, 0.75, 1 - 2 ** -53]])
    worst = 0.0
    for u in us:
        s, c = ob.sincos2pi(u)
        worst = max(worst, abs(s - math.sin(2 * math.pi * u)), abs(c - math.cos(2 * math.pi * u)))
        assert abs(s * s + c * c - 1.0) < 4e-16
    assert worst < 1e-15
    assert ob.sincos2pi(0.0) == (0.0, 1.0)


def test_quotient_below_one_iff_dividend_below_divisor():
    """The flight kernel decides 'face crossed this step' (l = num/den < 1,
    l clamped at 0) without dividing: num < den for den > 0, num > den for
    den < 0.  For IEEE round-to-nearest doubles this is exact; checked here on
    random pairs, on every adjacent pair (pred(b), b) and (b, succ(b)), and
    across binade edges, where a rounded quotient next to 1 could misround."""
    rng = np.random.default_rng(5)
    b = np.concatenate([rng.uniform(1e-9, 1e-2, 200000), 2.0 ** rng.integers(-40, 10, 20000),
                        np.nextafter(2.0 ** rng.integers(-40, 10, 20000), 0.0)])
    near = [np.nextafter(b, 0.0), b, np.nextafter(b, np.inf), np.nextafter(np.nextafter(b, 0.0), 0.0)]
    cands = near + [b * rng.uniform(0.0, 2.0, b.size), -b * rng.uniform(0.0, 1e-3, b.size)]
    for num in cands:
        for sgn in (1.0, -1.0):
            n_, d_ = sgn * num, sgn * b
            l = np.maximum(n_ / d_, 0.0)
            fast = (n_ < d_) if sgn > 0 else (n_ > d_)
            assert np.array_equal(l < 1.0, fast)


def test_reciprocal_division_is_exact():
    """The kernels split a cell index c = i + nx*(j + ny*k) with q = trunc(n*r),
    r = fl(1/d), plus one correction step (div_exact): exact integer division
    for 0 <= n < 2^31, 0 < d < 2^31, checked on random and boundary values
    (multiples of d and their neighbours) for small and large divisors."""
    rng = np.random.default_rng(3)
    for trial in range(60):
        d = int(rng.integers(1, 2**31)) if trial % 2 else int(rng.integers(1, 5000))
        mult = np.arange(0, 2**31 // d, max(1, (2**31 // d) // 20000), dtype=np.int64)[:20000] * d
        n = np.concatenate([rng.integers(0, 2**31, 20000), mult])
        n = np.concatenate([n, np.maximum(n - 1, 0), np.minimum(n + 1, 2**31 - 1)]).astype(np.int64)
        q = np.trunc(n.astype(np.float64) * (1.0 / d)).astype(np.int64)
        m = n - q * d
        q = np.where(m < 0, q - 1, np.where(m >= d, q + 1, q))
        assert np.array_equal(q, n // d)
