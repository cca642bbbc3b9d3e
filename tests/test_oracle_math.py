"""Pins of the oracle's single-precision transcendentals (Cephes-style sin, cos, exp, log, pow).

The shaders call sin/cos on [0, 2 pi] (photontrace.fs:437-439, 571-573; eyeraytrace.fs:417-418),
exp on [-inf, 0] and pow on (0, 1] (re.fs:19-42, glossy_reflect).  Each routine is compared with
the float64 value rounded to float32, over dense grids of those ranges and beyond, within a
bound in units in the last place; special values are checked exactly.
"""
import math

import numpy as np

import oracle


def ulp_err(got, ref):
    """|got - ref| in ulps of ref's float32 neighbourhood."""
    ref32 = np.float32(ref)
    spacing = abs(float(np.spacing(ref32))) if ref32 != 0 else float(np.finfo(np.float32).tiny)
    return abs(float(got) - ref) / spacing


def _check(fn, ref_fn, xs, bound):
    worst = 0.0
    for x in xs:
        x32 = float(np.float32(x))
        worst = max(worst, ulp_err(fn(x32), ref_fn(x32)))
    assert worst <= bound, worst
    return worst


def test_sin_cos_on_shader_range():
    xs = np.linspace(0.0, 6.283184, 20001)
    _check(oracle.sinf, math.sin, xs, 2.0)
    _check(oracle.cosf, math.cos, xs, 2.0)
    _check(oracle.sinf, math.sin, -xs[::7], 2.0)
    _check(oracle.cosf, math.cos, -xs[::7], 2.0)


def test_sin_cos_absolute_error_near_zeros():
    # near multiples of pi the relative error is unbounded for any float routine; the absolute
    # error stays at the level of the argument's rounding
    for k in range(1, 3):
        x = np.float32(k * math.pi)
        assert abs(float(oracle.sinf(x)) - math.sin(float(x))) < 2e-7
        x = np.float32((k - 0.5) * math.pi)
        assert abs(float(oracle.cosf(x)) - math.cos(float(x))) < 2e-7


def test_exp_range():
    xs = np.concatenate([np.linspace(-87.0, 88.0, 20001), -np.logspace(-8, 2, 2001)])
    _check(oracle.expf, math.exp, xs, 2.0)


def test_log_range():
    xs = np.concatenate([np.logspace(-37, 38, 20001), np.linspace(0.5, 2.0, 5001)])
    _check(oracle.logf, math.log, [x for x in xs if abs(x - 1.0) > 1e-3], 2.0)


def test_pow_on_unit_interval():
    # pow = exp(y log x): the rounding of y log x is amplified by |y ln x|, so the bound is
    # 3 + 1.5 |y ln x| ulp (4.7 ulp on the sRGB curve of re.fs, where |y ln x| < 2.5)
    worst = 0.0
    for x in np.linspace(1e-4, 1.0, 2001):
        for y in (0.416666666667, 0.1, 0.5, 0.9, 2.71828, 3.0):
            x32, y32 = float(np.float32(x)), float(np.float32(y))
            bound = 3.0 + 1.5 * abs(y32 * math.log(x32))
            err = ulp_err(oracle.powf(x32, y32), x32 ** y32)
            assert err <= bound, (x32, y32, err)
            worst = max(worst, err / bound)
    assert worst <= 1.0


def test_special_values():
    assert oracle.sinf(0.0) == 0.0
    assert oracle.cosf(0.0) == 1.0
    assert oracle.expf(0.0) == 1.0
    assert oracle.expf(-1000.0) == 0.0
    assert np.isinf(oracle.expf(100.0))
    assert oracle.logf(1.0) == 0.0
    assert np.isneginf(oracle.logf(0.0))
    assert oracle.powf(0.0, 0.4) == 0.0
    assert oracle.powf(0.3, 0.0) == 1.0
    assert oracle.powf(1.0, 7.0) == 1.0
    # the denormal end of log
    assert ulp_err(oracle.logf(1e-40), math.log(float(np.float32(1e-40)))) <= 2.0
