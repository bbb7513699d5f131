"""Pins of the oracle's random numbers (rng_par_zig.f90, TrackSubCell.F90).

The reference ships no RNG vectors, so the restatement is pinned by what the
mathematics fixes:
* the jump of rng_par_zig.f90:284-307 must equal 2^64 steps of the
  generator's own transition, computed independently as T^(2^64) over
  GF(2) by 64 squarings -- a wrong shift triple (24,16,37) or jump constant
  fails this;
* int64 -> real64 (rng_par_zig.f90:67-74) must give exactly (x >> 12)/2^52;
* the Irwin-Hall normal (TrackSubCell.F90:3631-3663) has mean 0, variance 1
  and support [-6, 6];
* stream p = jump^p(base) (rng_par_zig.f90:183-191) is consistent under
  offsets.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

BASE = (0x1DADBEEFBAADD0D0, 0x5BADD0D0DEADBEEF)


def _state(a, b):
    return (C.c_uint64 * 2)(a, b)


def _apply(fn, vec):
    s = _state(int(vec[0]), int(vec[1]))
    fn(s)
    return int(s[0]), int(s[1])


def _matrix(fn):
    """128x128 GF(2) matrix (column b = image of basis vector b)."""
    M = np.zeros((128, 128), dtype=np.int64)
    for b in range(128):
        v = (1 << b, 0) if b < 64 else (0, 1 << (b - 64))
        a0, a1 = _apply(fn, v)
        for r in range(64):
            M[r, b] = (a0 >> r) & 1
            M[64 + r, b] = (a1 >> r) & 1
    return M


def test_jump_equals_two_to_64_steps():
    L = O.lib()
    T = _matrix(lambda s: L.orc_xoroshiro_next(s))
    J = _matrix(lambda s: L.orc_rng_jump(s))
    P = T.copy()
    for _ in range(64):
        P = (P @ P) % 2
    assert np.array_equal(P, J)
    # and it is not trivially the identity or a single step
    assert not np.array_equal(J, np.eye(128, dtype=np.int64))
    assert not np.array_equal(J, T)


def test_uniform_is_mantissa_fill():
    L = O.lib()
    s1 = _state(*BASE)
    s2 = _state(*BASE)
    for _ in range(2000):
        x = L.orc_xoroshiro_next(s1)
        u = L.orc_uniform(s2)
        assert 0.0 <= u < 1.0
        assert u * 2.0 ** 52 == float(x >> 12)


def test_xoroshiro_first_outputs_are_sums():
    # output is s0 + s1 (mod 2^64) of the state before the step
    L = O.lib()
    s = _state(*BASE)
    a, b = int(s[0]), int(s[1])
    x = L.orc_xoroshiro_next(s)
    assert x == (a + b) % (1 << 64)


def test_normal12_moments_and_support():
    L = O.lib()
    s = _state(*BASE)
    r = np.array([L.orc_normal12(s) for _ in range(200000)])
    assert abs(r.mean()) < 0.01
    assert abs(r.var() - 1.0) < 0.01
    assert r.min() >= -6.0 and r.max() <= 6.0
    # Irwin-Hall kurtosis is 3 - 1.2/12 = 2.9
    k = ((r - r.mean()) ** 4).mean() / r.var() ** 2
    assert abs(k - 2.9) < 0.03


def test_normal12_is_sum_of_twelve_uniforms():
    L = O.lib()
    s1 = _state(*BASE)
    s2 = _state(*BASE)
    for _ in range(200):
        acc = 0.0
        for _ in range(12):
            acc = acc + L.orc_uniform(s1)
        assert acc - 6.0 == L.orc_normal12(s2)


def test_streams_offsets_consistent():
    a = O.rng_streams(12, first=0)
    b = O.rng_streams(5, first=7)
    assert np.array_equal(a[14:24], b)
    assert a[0] == np.uint64(BASE[0]) and a[1] == np.uint64(BASE[1])
