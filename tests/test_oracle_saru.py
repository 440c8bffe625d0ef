"""Pins of the oracle's Saru hash (force/pair_dpd.cu:212-236).

Two independent restatements (C oracle, numpy uint32 in tests/saru_numpy.py)
must agree bit for bit with each other and with the committed table; the
uniform conversions must respect the reference's FP32 rounding of s*2^-32 in
both builds; the stream must be uniform on [-sqrt3, sqrt3] with unit variance
and decorrelated across steps and tags.
"""
import os

import numpy as np
import pytest

from oracle import pyoracle
from saru_numpy import saru_bits, uniform_f32, uniform_f64

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "saru_table.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def test_table_matches_oracle(gold):
    for sd in np.unique(gold["seed"]):
        m = gold["seed"] == sd
        s, r32, r64 = pyoracle.saru_batch(gold["tag1"][m], gold["tag2"][m], int(sd), gold["step"][m])
        assert np.array_equal(s, gold["bits"][m])
        assert np.array_equal(r32.view(np.uint32), gold["r32"][m].view(np.uint32))
        assert np.array_equal(r64.view(np.uint64), gold["r64"][m].view(np.uint64))


def test_table_matches_numpy_restatement(gold):
    m = gold["seed"] == 42
    s = saru_bits(gold["tag1"][m], gold["tag2"][m], 42, gold["step"][m])
    assert np.array_equal(s, gold["bits"][m])
    assert np.array_equal(uniform_f32(s).view(np.uint32), gold["r32"][m].view(np.uint32))


def test_bulk_agreement_random_tuples():
    rng = np.random.Generator(np.random.PCG64(7))
    n = 100_000
    t1 = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    t2 = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    st = rng.integers(-(2 ** 31), 2 ** 31, n, dtype=np.int64).astype(np.int32)
    s_np = saru_bits(t1, t2, 123456, st)
    s_c, _, _ = pyoracle.saru_batch(t1, t2, 123456, st)
    assert np.array_equal(s_np, s_c)


def test_uniform_conversion_rules():
    # the reference multiplies the uint32 by a float literal 2^-32: FP32 rounding of s
    s = np.array([0, 1, 0xFFFFFFFF, 0x80000000, 0x7FFFFFFF, 0xFFFFFF7F, 0xFFFFFF80], dtype=np.uint32)
    r64 = uniform_f64(s)
    assert r64[0] == -1.7320508075688772935274463
    # 0xFFFFFFFF rounds to 2^32 in FP32, so u = 1 exactly and r = +sqrt(3)
    assert r64[2] == 1.7320508075688772935274463
    # 0xFFFFFF7F rounds down to 2^32-256 in FP32 (round to nearest), 0xFFFFFF80 ties to even -> 2^32
    assert r64[5] < r64[6] == r64[2]
    r32 = uniform_f32(s)
    assert np.all(np.abs(r32 - r64) < 2e-7)


def test_distribution_and_decorrelation():
    n = 200_000
    ids = np.arange(1, n + 1, dtype=np.uint32)
    s0 = saru_bits(ids, ids + 1, 42, np.zeros(n, np.int32))
    s1 = saru_bits(ids, ids + 1, 42, np.ones(n, np.int32))
    r0, r1 = uniform_f64(s0), uniform_f64(s1)
    sig = 1.0 / np.sqrt(n)
    assert abs(r0.mean()) < 5 * sig
    assert abs(r0.var() - 1.0) < 5 * 0.9 * sig  # var of U(-sqrt3,sqrt3)^2 is 0.8
    assert r0.min() >= -np.sqrt(3) - 1e-12 and r0.max() <= np.sqrt(3) + 1e-12
    assert abs(np.corrcoef(r0, r1)[0, 1]) < 5 * sig
    # neighbouring tag pairs are decorrelated
    assert abs(np.corrcoef(r0[:-1], r0[1:])[0, 1]) < 5 * sig
    # flat histogram
    h, _ = np.histogram(r0, bins=20, range=(-np.sqrt(3), np.sqrt(3)))
    assert np.all(np.abs(h - n / 20) < 6 * np.sqrt(n / 20))


def test_seed_and_step_enter_only_as_their_sum():
    # seed3 = seed + timestep (pair_dpd.cu:213): (seed, step) pairs with equal sums coincide
    t1 = np.arange(1, 1001, dtype=np.uint32)
    a = saru_bits(t1, t1 + 5, 100, np.full(1000, 7, np.int32))
    b = saru_bits(t1, t1 + 5, 107, np.zeros(1000, np.int32))
    c = saru_bits(t1, t1 + 5, 100, np.full(1000, 8, np.int32))
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    # tag order matters, which is why the call sites sort the ids
    d = saru_bits(t1 + 5, t1, 100, np.full(1000, 7, np.int32))
    assert not np.array_equal(a, d)
