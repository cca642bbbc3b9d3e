"""Pins of the oracle's random number generation.

- XORShift::m_frand (glslppm.cpp:77-91) against an independent integer restatement of Marsaglia's
  xorshift128 (J. Statistical Software 8(14), 2003) and its float mapping u32 * 2^-32.
- RandomizeTextures (glslppm.cpp:113-140): texel layout of the seed grids.
- GPURnd (eyeraytrace.fs:326-347): each lane is Schrage's factorisation of the multiplicative
  congruential step n -> a n mod m with m = a q + r, checked against exact integer arithmetic,
  and the output against the double-precision fract(n0/m0 - n1/m1 + n2/m2 - n3/m3).
"""
import numpy as np
import pytest

import oracle

SEED = [123456789, 362436069, 521288629, 88675123]
A = [3423, 2646, 1707, 1999]
Q = [1225, 1585, 2457, 2098]
R = [1112, 367, 92, 265]
M = [4194287, 4194277, 4194191, 4194167]


def xorshift128(state):
    x, y, z, w = state
    t = (x ^ (x << 11)) & 0xFFFFFFFF
    x, y, z = y, z, w
    w = (w ^ (w >> 19) ^ (t ^ (t >> 8))) & 0xFFFFFFFF
    return [x, y, z, w], w


def test_schrage_constants_factor_each_modulus():
    for a, q, r, m in zip(A, Q, R, M):
        assert a * q + r == m and r < q  # Schrage's condition keeps every step inside int range


def test_xorshift_sequence_and_float_mapping():
    st = list(SEED)
    ost = list(SEED)
    for _ in range(20000):
        st, u = xorshift128(st)
        v, ost = oracle.xorshift_frand(ost)
        assert ost == st
        assert v == np.float32(np.float32(u) * np.float32(2.0 ** -32))


def test_first_xorshift_outputs_known_answer():
    # Marsaglia's xorshift128 from the seed above: the first outputs, computed independently
    st = list(SEED)
    outs = []
    for _ in range(3):
        st, u = xorshift128(st)
        outs.append(u)
    assert outs == [3701687786, 458299110, 2500872618]


def test_seed_grid_layout_last_argument_first():
    grid, st = oracle.seed_grid(SEED, 64)
    ref = list(SEED)
    for i in range(64):
        draws = []
        for _ in range(4):
            ref, u = xorshift128(ref)
            draws.append(np.float32(np.float32(u) * np.float32(2.0 ** -32)) * np.float32(4194304.0))
        # g++ evaluates TVector4(a, b, c, d) arguments from d to a: w gets the first draw
        assert list(grid[i]) == [draws[3], draws[2], draws[1], draws[0]]
    assert st == ref


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_gpurnd_integer_states_follow_mlcg(seed):
    rng = np.random.default_rng(seed)
    n = rng.integers(1, np.array(M), size=4)
    state = n.astype(np.float32)
    exact = [int(v) for v in n]
    for step in range(3000):
        out = oracle.gpurnd(state)
        exact = [(a * v) % m for a, v, m in zip(A, exact, M)]
        assert [int(v) for v in state] == exact, step
        assert np.all(state == np.float32(exact))
        ref = sum(s * v / m for s, v, m in zip((1, -1, 1, -1), exact, M)) % 1.0
        # error bound of the float evaluation: four quotients n * fl(1/m) below 1 with relative
        # error 2^-23, three additions of magnitude below 2 rounding by 2^-24 each: 11 * 2^-24
        err = abs(float(out) - ref)
        assert min(err, abs(err - 1.0)) <= 11 * 2.0 ** -24


def test_gpurnd_zero_lane_goes_to_half_modulus():
    # sign(0) = 0 in GLSL: p = 0 maps to m/2 instead of 0
    state = np.zeros(4, np.float32)
    oracle.gpurnd(state)
    assert list(state) == [np.float32(m / 2) for m in M]


def test_gpurnd_fractional_seeds_stay_congruent_up_to_rounding():
    # real seeds are float(u32) * 2^-10, with fractional bits; the product a (n - beta q) then
    # rounds (magnitude below 2^23, so by at most 1/2) and the rest of the step is exact:
    # c = a p (mod m) within 1/2, and c stays in [0, m]
    grid, _ = oracle.seed_grid(SEED, 256)
    from fractions import Fraction
    fractional = 0
    for row in grid[:64]:
        state = row.copy()
        prev = [Fraction(float(v)) for v in state]
        for _ in range(50):
            oracle.gpurnd(state)
            cur = [Fraction(float(v)) for v in state]
            for a, m, p, c in zip(A, M, prev, cur):
                d = a * p - c
                k = round(d / m)
                assert abs(d - k * m) <= Fraction(1, 2)
                assert 0 <= c <= m
                fractional += p.denominator != 1
            prev = cur
    assert fractional > 0


def test_gpurnd_output_is_uniform():
    state = np.array([1234567.0, 2345678.0, 3456789.0, 456789.0], np.float32)
    vals = np.array([oracle.gpurnd(state) for _ in range(20000)], np.float64)
    assert np.all((vals >= 0) & (vals < 1))
    hist, _ = np.histogram(vals, bins=20, range=(0, 1))
    expected = len(vals) / 20
    chi2 = float(((hist - expected) ** 2 / expected).sum())
    assert chi2 < 45.3  # 19 degrees of freedom, p = 0.999
