"""Pins the oracle's Philox4x32-10, its u64 stream layout, the rand-0.5 style
u64 -> f64 maps and the shared exp() against published vectors, rocRAND's
source header, exact integer arithmetic and the C library's exp."""
import math
import os
import struct
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle_ffi as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module", autouse=True)
def _lib():
    O.lib()


# Random123 kat_vectors for philox4x32_10 (Salmon et al., SC'11).
R123_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", R123_KAT)
def test_philox_random123_kat(ctr, key, want):
    assert O.philox(ctr, key) == want


def test_philox_matches_rocrand_header_vectors():
    rows = []
    with open(os.path.join(GOLDEN, "philox_rocrand.txt")) as f:
        for line in f:
            if line.startswith("#"):
                continue
            v = [int(t, 16) for t in line.split()]
            rows.append(v)
    assert len(rows) == 64
    for v in rows:
        assert O.philox(tuple(v[0:4]), tuple(v[4:6])) == tuple(v[6:10])


def test_stream_layout_low_word_first():
    L = O.lib()
    seed, sys_, w, ep = 0x0123456789ABCDEF, 3, 77, 5
    for i in [0, 1, 2, 3, 1000, 1001, (1 << 33) + 1]:
        blk = i >> 1
        ctr = (blk & 0xFFFFFFFF, blk >> 32, w, ep)
        key = (seed & 0xFFFFFFFF, (seed >> 32) ^ sys_)
        words = O.philox(ctr, key)
        j = i & 1
        want = words[2 * j] | (words[2 * j + 1] << 32)
        assert L.orc_stream_u64(seed, sys_, w, ep, i) == want


def _bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def _from_bits(b):
    return struct.unpack("<d", struct.pack("<Q", b))[0]


def test_standard_map_exact():
    L = O.lib()
    rng = np.random.default_rng(1)
    xs = [0, 1, (1 << 11) - 1, 1 << 11, (1 << 64) - 1, 1 << 63] + [int(v) for v in rng.integers(0, 2**63, 200, dtype=np.int64)]
    for x in xs:
        got = L.orc_u64_standard(x)
        want = Fraction(x >> 11, 1 << 53)
        assert Fraction(got) == want  # exact: 53-bit integer times a power of two
        assert 0.0 <= got < 1.0
    assert L.orc_u64_standard((1 << 64) - 1) == 1.0 - 2.0**-53


def test_range_map_matches_two_rounding_recipe():
    L = O.lib()
    rng = np.random.default_rng(2)
    for lo, hi in [(-1.0, 1.0), (-0.25, 0.25), (-0.5, 0.5), (0.5, 7.0)]:
        scale = hi - lo
        offset = lo - scale
        for x in [0, (1 << 64) - 1] + [int(v) for v in rng.integers(0, 2**63, 300, dtype=np.int64)]:
            v = _from_bits(0x3FF0000000000000 | (x >> 12))
            want = np.float64(v) * np.float64(scale)  # rounded product
            want = float(want + np.float64(offset))    # rounded sum
            got = L.orc_u64_range(x, lo, hi)
            assert got == want
            assert lo <= got < hi


def _ulp_diff(a, b):
    ia, ib = _bits(a), _bits(b)
    return abs(ia - ib)


def test_exp_within_one_ulp_of_libm():
    L = O.lib()
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-50, 50, 20000), rng.uniform(-1, 1, 20000),
                         rng.uniform(-700, 700, 5000), rng.uniform(-1e-8, 1e-8, 1000)])
    worst = 0
    for x in xs:
        worst = max(worst, _ulp_diff(L.orc_exp(float(x)), math.exp(float(x))))
    assert worst <= 1


def test_exp_special_values():
    L = O.lib()
    assert L.orc_exp(0.0) == 1.0
    assert L.orc_exp(-0.0) == 1.0
    assert L.orc_exp(1.0) == math.exp(1.0) or _ulp_diff(L.orc_exp(1.0), math.e) <= 1
    assert math.isinf(L.orc_exp(710.0))
    assert L.orc_exp(-746.0) == 0.0
    assert math.isnan(L.orc_exp(float("nan")))
    assert math.isinf(L.orc_exp(float("inf")))
    assert L.orc_exp(float("-inf")) == 0.0
    # subnormal range stays monotone and within a few ulp of libm
    for x in [-708.5, -720.0, -740.0, -744.0]:
        assert _ulp_diff(L.orc_exp(x), math.exp(x)) <= 1


def test_norm3_is_left_to_right_sum_of_squares():
    L = O.lib()
    rng = np.random.default_rng(4)
    for _ in range(500):
        x, y, z = rng.normal(size=3) * 10.0 ** rng.integers(-3, 3)
        s = np.float64(x) * x
        s = s + np.float64(y) * y
        s = s + np.float64(z) * z
        assert L.orc_norm3(x, y, z) == float(np.sqrt(s))


def _ulps(a, b):
    return abs(a - b) / math.ulp(abs(b)) if b != 0 else abs(a)


def test_log_within_two_ulp_of_libm():
    rng = np.random.default_rng(11)
    xs = list(rng.uniform(0.0, 1.0, 20000)) + [2.0 ** -k for k in range(54)]
    xs += [1.0 - 2.0 ** -53, 0.7071067811865476, 1.4142135623730951, 3.0, 1e300, 1e-300]
    worst = 0.0
    for x in xs:
        r = math.log(x)
        v = O.log(x)
        worst = max(worst, _ulps(v, r) if r != 0 else abs(v))
    assert worst <= 2.0
    assert O.log(1.0) == 0.0


def test_sincospi_against_libm_and_exact_points():
    rng = np.random.default_rng(12)
    for x in rng.uniform(0.0, 2.0, 20000):
        s, c = O.sincospi(x)
        # pi*x rounds once in libm's argument; allow that and a few ulp
        assert abs(s - math.sin(math.pi * x)) <= 2e-15
        assert abs(c - math.cos(math.pi * x)) <= 2e-15
    exact = {0.0: (0.0, 1.0), 0.5: (1.0, 0.0), 1.0: (0.0, -1.0), 1.5: (-1.0, 0.0)}
    for x, (s0, c0) in exact.items():
        s, c = O.sincospi(x)
        assert abs(s - s0) <= 1e-300 and abs(c - c0) <= 1e-300, x
    s, c = O.sincospi(0.25)
    assert abs(s - math.sqrt(0.5)) <= 1.2e-16 and abs(c - math.sqrt(0.5)) <= 1.2e-16


def test_box_muller_is_standard_normal_and_follows_its_recipe():
    from scipy import stats
    n = 100000
    a = [O.lib().orc_stream_u64(5, 0, w, 0, 0) for w in range(n)]
    b = [O.lib().orc_stream_u64(5, 0, w, 0, 1) for w in range(n)]
    z = np.array([O.box_muller(x, y) for x, y in zip(a, b)])
    for col in (z[:, 0], z[:, 1], z.ravel()):
        assert stats.kstest(col, "norm").pvalue > 1e-3
    # pairs are uncorrelated
    assert abs(np.corrcoef(z[:, 0], z[:, 1])[0, 1]) < 0.02
    # recipe: r = sqrt(-2 ln(1 - U(a))), angle 2 pi U(b)
    for x, y in list(zip(a, b))[:200]:
        u1 = 1.0 - O.lib().orc_u64_standard(x)
        u2 = O.lib().orc_u64_standard(y)
        r = math.sqrt(-2.0 * math.log(u1))
        z0, z1 = O.box_muller(x, y)
        assert abs(z0 - r * math.cos(2 * math.pi * u2)) <= 1e-13 * max(1.0, r)
        assert abs(z1 - r * math.sin(2 * math.pi * u2)) <= 1e-13 * max(1.0, r)
    # u1 is never 0: the largest draw still gives a finite radius
    z0, z1 = O.box_muller(0xFFFFFFFFFFFFFFFF, 0)
    assert math.isfinite(z0) and z0 > 8.0
