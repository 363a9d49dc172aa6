"""Device primitives of the Gaussian proposals and of the walker MLP, against
the CPU oracle (bit for bit) or an exact host product: ln, sin(pi x),
cos(pi x), Box-Muller normals, bf16 rounding and the lane placement of one
v_mfma_f32_16x16x16_bf16."""
import numpy as np
import pytest

import mole_amd as M
from oracle import oracle_ffi as O

pytestmark = pytest.mark.gpu


def bits(a):
    return np.asarray(a, dtype=np.float64).view(np.uint64)


def test_device_log_bitwise_matches_oracle():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(0, 1, 40000), 2.0 ** -np.arange(54.0), [1.0 - 2.0**-53, 3.0, 1e300]])
    got = M.debug_math(M.DEBUG_LOG, x)
    want = np.array([O.log(v) for v in x])
    assert np.array_equal(bits(got), bits(want))


def test_device_sincospi_bitwise_matches_oracle():
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.uniform(0, 2, 40000), np.arange(0, 2, 0.125)])
    s = M.debug_math(M.DEBUG_SINPI, x)
    c = M.debug_math(M.DEBUG_COSPI, x)
    want = np.array([O.sincospi(v) for v in x])
    assert np.array_equal(bits(s), bits(want[:, 0]))
    assert np.array_equal(bits(c), bits(want[:, 1]))


def test_device_box_muller_bitwise_matches_oracle():
    rng = np.random.default_rng(5)
    a = rng.integers(0, 2**63, 20000, dtype=np.int64).astype(np.uint64) * np.uint64(2)
    b = rng.integers(0, 2**63, 20000, dtype=np.int64).astype(np.uint64) * np.uint64(2) + np.uint64(1)
    a = np.concatenate([a, np.array([2**64 - 1, 0], dtype=np.uint64)])
    b = np.concatenate([b, np.array([0, 2**64 - 1], dtype=np.uint64)])
    z0 = M.debug_math(M.DEBUG_BOXMULLER0, a.view(np.float64), b.view(np.float64))
    z1 = M.debug_math(M.DEBUG_BOXMULLER1, a.view(np.float64), b.view(np.float64))
    want = np.array([O.box_muller(int(x), int(y)) for x, y in zip(a, b)])
    assert np.array_equal(bits(z0), bits(want[:, 0]))
    assert np.array_equal(bits(z1), bits(want[:, 1]))


def bf16_round(x):
    """Host bf16 round-to-nearest-even of float32 values."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def test_device_bf16_rounding():
    rng = np.random.default_rng(6)
    x = np.concatenate([rng.normal(size=20000) * 10.0 ** rng.uniform(-8, 8, 20000),
                        [0.0, 1.0, 1.0 + 2**-8, 1.0 + 2**-9, 1.0 + 3 * 2**-9, -2.5]]).astype(np.float32)
    got = M.debug_math(M.DEBUG_BF16, x.astype(np.float64))
    assert np.array_equal(got.astype(np.float32), bf16_round(x))
    # ties to even: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7 -> 1;
    # 1 + 3 2^-9 is above the midpoint -> 1 + 2^-7
    assert got[-4] == 1.0 and got[-2] == 1.0 + 2**-7


def test_mfma_bf16_lane_placement():
    rng = np.random.default_rng(7)
    # small integers are exact in bf16 and their products sum exactly in FP32
    A = rng.integers(-8, 9, (16, 16)).astype(np.float32)
    B = rng.integers(-8, 9, (16, 16)).astype(np.float32)
    D = M.debug_mfma_bf16(A, B)
    assert np.array_equal(D, A @ B)
    # distinct row/column structure: a transposed placement would fail here
    A = np.arange(256, dtype=np.float32).reshape(16, 16) % 7 - 3
    B = (np.arange(256, dtype=np.float32).reshape(16, 16).T % 5) - 2
    assert np.array_equal(M.debug_mfma_bf16(A, B), A @ B)


def test_mfma_bf16_k32_lane_placement():
    rng = np.random.default_rng(8)
    A = rng.integers(-8, 9, (16, 32)).astype(np.float32)
    B = rng.integers(-8, 9, (32, 16)).astype(np.float32)
    assert np.array_equal(M.debug_mfma_bf16_k32(A, B), A @ B)
    A = (np.arange(512, dtype=np.float32).reshape(16, 32) % 7) - 3
    B = (np.arange(512, dtype=np.float32).reshape(32, 16) % 5) - 2
    assert np.array_equal(M.debug_mfma_bf16_k32(A, B), A @ B)
