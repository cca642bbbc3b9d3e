"""The binary16 node boxes of the photon and eye walks: bounds rounded outward must make the
reference's slab test (eyeraytrace.fs:92-116, photontrace.fs:148-172) pass whenever it passes on
the exact fp32 box, with an entry distance no larger, so that the walk only ever enters more boxes
than the reference and the exact leaf re-test decides.  Checked on random and adversarial rays
(zero direction components, whose slabs go NaN, origins on the box planes) and boxes (bounds that
straddle zero, tiny, beyond the binary16 range), and on the node boxes of a generated scene.
"""
import numpy as np

import glslppm_amd as g

f32 = np.float32


def outward(x, down):
    """Largest binary16 <= x (down) or smallest >= x (up), keeping the sign of x and never a
    subnormal (the packing rule of the node records)."""
    x = np.asarray(x, f32)
    with np.errstate(over="ignore"):
        h = x.astype(np.float16)
        for _ in range(4):
            v = h.astype(f32)
            bad = (v > x) if down else (v < x)
            if not bad.any():
                break
            step = np.float16(-np.inf) if down else np.float16(np.inf)
            h = np.where(bad, np.nextafter(h, step), h)
        # subnormals step further outward: to a signed zero or the smallest normal (6.1e-5)
        v = h.astype(f32)
        sub = (v != 0) & (np.abs(v) < f32(2.0 ** -14))
        if down:
            v = np.where(sub, np.where(v < 0, f32(-(2.0 ** -14)), f32(0.0)), v)
        else:
            v = np.where(sub, np.where(v < 0, f32(-0.0), f32(2.0 ** -14)), v)
    return v.astype(f32)


def slab(rcp, t0, lo, hi, ht):
    """tmin, and whether the box passes, in the kernels' float32 order with NaN-rejecting
    min/max (fmin/fmax)."""
    with np.errstate(all="ignore"):
        l = (rcp * lo).astype(f32) + t0
        h = (rcp * hi).astype(f32) + t0
        mn, mx = np.fmin(l, h), np.fmax(l, h)
        tmin = np.fmax(np.fmax(mn[:, 0], mn[:, 1]), mn[:, 2])
        tmax = np.fmin(np.fmin(mx[:, 0], mx[:, 1]), mx[:, 2])
        ok = (tmin <= tmax) & (tmin <= ht) & (tmax >= 0)
    return tmin, ok


def rays(rs, n):
    d = rs.normal(size=(n, 3)).astype(f32)
    o = rs.uniform(-2, 2, (n, 3)).astype(f32)
    # zero components of either sign, and origins on the box planes used below
    z = rs.random((n, 3)) < 0.15
    d[z] = np.where(rs.random(z.sum()) < 0.5, f32(0.0), f32(-0.0))
    o[rs.random((n, 3)) < 0.1] = f32(0.0)
    with np.errstate(all="ignore"):
        rcp = (f32(1.0) / d).astype(f32)
        t0 = (-rcp * o).astype(f32)
    return rcp, t0, o, d


def check(rcp, t0, lo, hi, ht):
    lo16, hi16 = outward(lo, True), outward(hi, False)
    assert np.all(lo16 <= lo) and np.all(hi16 >= hi)
    assert np.all(np.signbit(lo16) <= np.signbit(lo) | (lo < 0))  # a bound >= +0 stays >= +0
    tmin_e, ok_e = slab(rcp, t0, lo, hi, ht)
    tmin_c, ok_c = slab(rcp, t0, lo16, hi16, ht)
    assert np.all(ok_c[ok_e]), int((ok_e & ~ok_c).sum())
    assert np.all(tmin_c[ok_e] <= tmin_e[ok_e])
    return int(ok_e.sum()), int(ok_c.sum())


def test_outward_rounding_brackets_and_keeps_sign():
    xs = np.array([0.0, -0.0, 1e-9, -1e-9, 6e-8, -6e-8, 0.1, -0.1, 1.0, 65504.0, 65520.0, 7e4, -7e4, 3.0e38],
                  f32)
    lo, hi = outward(xs, True), outward(xs, False)
    assert np.all(lo <= xs) and np.all(hi >= xs)
    # positive values never round down below +0, negative ones never up above -0
    assert np.all(lo[xs > 0] >= 0) and np.all(hi[xs < 0] <= 0)
    assert lo[0] == 0 and not np.signbit(lo[0]) and np.signbit(hi[1])
    assert lo[4] == 0 and hi[4] == 2.0 ** -14 and lo[5] == -(2.0 ** -14) and hi[5] == 0 and np.signbit(hi[5])
    assert lo[11] == 65504.0 and hi[11] == np.inf and lo[12] == -np.inf and hi[12] == -65504.0


def test_rounded_box_passes_whenever_exact_box_passes_random():
    rs = np.random.default_rng(3)
    n = 200000
    rcp, t0, o, d = rays(rs, n)
    c = rs.uniform(-2, 2, (n, 3)).astype(f32)
    on_ray = rs.random(n) < 0.5  # half the boxes around a point of their ray
    c[on_ray] = (o + rs.uniform(-0.5, 3, (n, 1)).astype(f32) * d)[on_ray]
    ext = np.exp(rs.uniform(np.log(1e-6), np.log(3.0), (n, 3))).astype(f32)
    lo = (c - ext).astype(f32)
    hi = (c + ext).astype(f32)
    # some bounds exactly zero, some straddling it by a hair, some beyond the binary16 range
    lo[rs.random((n, 3)) < 0.05] = f32(0.0)
    hi[rs.random((n, 3)) < 0.05] = f32(-0.0)
    hi = np.maximum(hi, lo)
    big = rs.random(n) < 0.02
    lo[big] *= f32(1e5)
    hi[big] *= f32(1e5)
    ht = np.where(rs.random(n) < 0.5, f32(1e30), rs.uniform(0, 3, n).astype(f32))
    n_e, n_c = check(rcp, t0, lo, hi, ht)
    assert n_e > 10000 and n_c >= n_e


def test_rounded_scene_nodes_pass_whenever_exact_nodes_pass():
    m = g.Mesh.generate(g.SCENE_SPHERES, seed=11, count=12, detail=3).build()
    b = m.arrays()["node_bounds"]
    rs = np.random.default_rng(5)
    idx = rs.integers(0, b.shape[0], 100000)
    rcp, t0, _, _ = rays(rs, idx.size)
    ht = np.where(rs.random(idx.size) < 0.5, f32(1e30), rs.uniform(0, 2, idx.size).astype(f32))
    n_e, n_c = check(rcp, t0, b[idx, :3].astype(f32), b[idx, 3:].astype(f32), ht)
    assert n_e > 100
    # rounding outward admits a few more boxes, not many more
    assert n_c <= 1.2 * n_e + 50


def test_nearest_rounding_would_break_the_walk():
    """The same check catches the plausible mistake of rounding to nearest."""
    rs = np.random.default_rng(8)
    n = 200000
    rcp, t0, o, d = rays(rs, n)
    c = (o + rs.uniform(0.1, 2, (n, 1)).astype(f32) * d).astype(f32)
    ext = np.exp(rs.uniform(np.log(1e-4), np.log(1e-1), (n, 3))).astype(f32)
    lo, hi = (c - ext).astype(f32), (c + ext).astype(f32)
    with np.errstate(all="ignore"):
        lo_n, hi_n = lo.astype(np.float16).astype(f32), hi.astype(np.float16).astype(f32)
    _, ok_e = slab(rcp, t0, lo, hi, f32(1e30))
    _, ok_n = slab(rcp, t0, lo_n, hi_n, f32(1e30))
    assert (ok_e & ~ok_n).sum() > 0
