"""Pins for the oracle's scalar building blocks: atan2f against libm,
Morton bit interleaving against a bit-by-bit construction, clz against
Python's bit_length."""
import numpy as np
import pytest

from harness import oracle


def ulp_err(got: np.ndarray, want64: np.ndarray) -> np.ndarray:
    want32 = want64.astype(np.float32)
    spacing = np.spacing(np.abs(want32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - want64) / spacing


def test_atan2f_within_3ulp_of_libm():
    """The reduction adds up to two roundings (t = min/max, pi/2 - r or pi - r):
    observed worst case 2.6 ulp; a wrong octant or sign would be ~1e7 ulp."""
    lib = oracle()
    rng = np.random.default_rng(1)
    y = (rng.standard_normal(40000) * 10 ** rng.uniform(-6, 6, 40000)).astype(np.float32)
    x = (rng.standard_normal(40000) * 10 ** rng.uniform(-6, 6, 40000)).astype(np.float32)
    got = np.array([lib.atan2f(float(a), float(b)) for a, b in zip(y, x)], np.float32)
    want = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    err = ulp_err(got, want)
    assert err.max() <= 3.0, f"max ulp error {err.max()}"


def test_atan2f_octants_and_special_cases():
    lib = oracle()
    pi = np.float32(np.pi)
    cases = [
        (0.0, 1.0, 0.0), (0.0, -1.0, np.pi), (1.0, 0.0, np.pi / 2), (-1.0, 0.0, -np.pi / 2),
        (1.0, 1.0, np.pi / 4), (1.0, -1.0, 3 * np.pi / 4), (-1.0, -1.0, -3 * np.pi / 4),
        (np.inf, np.inf, np.pi / 4), (np.inf, 1.0, np.pi / 2), (1.0, -np.inf, np.pi), (-1.0, np.inf, -0.0),
    ]
    for y, x, want in cases:
        got = lib.atan2f(y, x)
        assert abs(got - want) <= 2 * np.spacing(np.float32(abs(want)) + np.float32(1e-30)), (y, x, got, want)
    # C99 signed-zero rules
    assert np.copysign(1, lib.atan2f(0.0, 0.0)) > 0
    assert np.copysign(1, lib.atan2f(-0.0, 0.0)) < 0
    assert lib.atan2f(0.0, -0.0) == pi
    assert lib.atan2f(-0.0, -0.0) == -pi
    assert np.isnan(lib.atan2f(float("nan"), 1.0))
    assert np.isnan(lib.atan2f(1.0, float("nan")))


def spread_bits(v: int) -> int:
    out = 0
    for b in range(10):
        out |= ((v >> b) & 1) << (3 * b)
    return out


def test_expand_bits_matches_bitwise_spread():
    lib = oracle()
    for v in range(1024):
        assert lib.expand_bits(v) == spread_bits(v)
    assert lib.expand_bits(0x3FF) == 0x09249249
    assert lib.expand_bits(0xFFFFF) == 0x09249249  # upper bits masked off


def test_morton3d_known_answers():
    lib = oracle()
    assert lib.morton3d(0.0, 0.0, 0.0) == 0
    assert lib.morton3d(1.0, 1.0, 1.0) == 0x3FFFFFFF  # 1.0*1024 clamps to 1023
    assert lib.morton3d(2.0, -1.0, 0.5) == (spread_bits(1023) << 2) | (spread_bits(0) << 1) | spread_bits(512)
    rng = np.random.default_rng(2)
    for _ in range(500):
        x, y, z = rng.uniform(-0.1, 1.1, 3).astype(np.float32)
        q = [int(min(max(np.float32(c) * np.float32(1024), 0), 1023)) for c in (x, y, z)]
        want = (spread_bits(q[0]) << 2) | (spread_bits(q[1]) << 1) | spread_bits(q[2])
        assert lib.morton3d(float(x), float(y), float(z)) == want


@pytest.mark.parametrize("v", [0, 1, 2, 3, 0x80000000, 0xFFFFFFFF, 0x00010000, 0x7FFFFFFF, 12345])
def test_clz32(v):
    assert oracle().clz32(v) == 32 - v.bit_length()
