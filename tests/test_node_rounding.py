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


def fma32(a, b, c):
    """float32(a * b + c) with one rounding (to nearest even) for float32 a, b, c: the product is
    exact in float64, the sum's error is recovered exactly (two-sum), and a float64 sum that
    falls on a float32 midpoint is resolved by that error's sign."""
    with np.errstate(all="ignore"):
        p = a.astype(np.float64) * b.astype(np.float64)
        c64 = c.astype(np.float64)
        s = p + c64
        bb = s - p
        e = (p - (s - bb)) + (c64 - bb)
        r = s.astype(f32)
        r64 = r.astype(np.float64)
        n = np.nextafter(r, np.where(s > r64, f32(np.inf), f32(-np.inf)).astype(f32))
        n64 = n.astype(np.float64)
        mid = np.isfinite(s) & (s != r64) & ((r64 + n64) * 0.5 == s) & (e != 0)
        toward_n = np.sign(e) == np.sign(n64 - s)
        return np.where(mid, np.where(toward_n, n, r), r).astype(f32)


def test_fma32_is_correctly_rounded():
    """Against exact rational arithmetic on cases built to land near float32 midpoints."""
    from fractions import Fraction
    rs = np.random.default_rng(21)
    n = 3000
    a = rs.uniform(-4, 4, n).astype(f32)
    b = outward(rs.uniform(-4, 4, n).astype(f32), True)
    c = (-(a.astype(np.float64) * b.astype(np.float64))).astype(f32)  # cancellation
    c[: n // 2] = rs.uniform(-1e4, 1e4, n // 2).astype(f32)
    got = fma32(a, b, c)
    for ai, bi, ci, gi in zip(a, b, c, got):
        x = Fraction(float(ai)) * Fraction(float(bi)) + Fraction(float(ci))
        lo = f32(float(x))
        cand = [lo, np.nextafter(lo, f32(np.inf)), np.nextafter(lo, f32(-np.inf))]
        dist = [abs(Fraction(float(v)) - x) for v in cand]
        best = min(dist)
        ties = [v for v, d in zip(cand, dist) if d == best]
        want = ties[0] if len(ties) == 1 else [v for v in ties if int(v.view(np.uint32)) % 2 == 0][0]
        assert gi == want, (ai, bi, ci, gi, want)


def sign_order_rays(rcp, t0):
    """The rays the four-wide record tests serve (sign_order_ok in kernels.hip)."""
    r = np.abs(rcp)
    return np.all((r >= f32(2.0 ** -100)) & (r <= f32(2.0 ** 100)), axis=1) & np.all(np.abs(t0) <= f32(2.0 ** 120), axis=1)


def slab_eps(lo, hi):
    """2^-22 times the largest bound magnitude capped at 2^17, at least 2^-126 (the scene constant
    the record packing computes)."""
    b = np.fmin(np.max(np.abs(np.concatenate([lo.ravel(), hi.ravel()]))), f32(2.0 ** 17))
    return f32(max(float(b) * 2.0 ** -22, 2.0 ** -126))


def slack(rcp, t0, eps):
    """e = |rcp| eps + 2^-22 |t0| + 2^-140 in two fp32 fmas, as slab_slack evaluates it."""
    inner = fma32(np.abs(t0), np.full_like(t0, f32(2.0 ** -22)), np.full_like(t0, f32(2.0 ** -140)))
    return fma32(np.abs(rcp), np.full_like(rcp, eps), inner)


def wide_slab(rcp, t0, lo16, hi16, ht, eps, with_slack=True):
    """The four-wide record test: near and far bounds by the reciprocal's sign, each slab value one
    fused multiply-add with the offset moved outward by the slack, for rays that pass
    sign_order_rays."""
    e = slack(rcp, t0, eps) if with_slack else np.zeros_like(t0)
    near = np.where(rcp < 0, hi16, lo16)
    far = np.where(rcp < 0, lo16, hi16)
    n = fma32(rcp, near, (t0 - e).astype(f32))
    f = fma32(rcp, far, (t0 + e).astype(f32))
    tmin = np.max(n, axis=1)
    tmax = np.min(f, axis=1)
    return tmin, (tmin <= tmax) & (tmin <= ht) & (tmax >= 0)


def record_cases(rs, n):
    rcp, t0, o, d = rays(rs, n)
    c = rs.uniform(-2, 2, (n, 3)).astype(f32)
    on_ray = rs.random(n) < 0.5
    c[on_ray] = (o + rs.uniform(-0.5, 3, (n, 1)).astype(f32) * d)[on_ray]
    ext = np.exp(rs.uniform(np.log(1e-6), np.log(3.0), (n, 3))).astype(f32)
    lo, hi = (c - ext).astype(f32), (c + ext).astype(f32)
    # bounds a binary16 holds exactly (the packed bound is the exact one: no rounding slack), on
    # the ray's origin plane, and zero
    snap = rs.random((n, 3)) < 0.5
    lo[snap] = np.floor(lo[snap] * 512) / 512
    hi[snap] = np.ceil(hi[snap] * 512) / 512
    plane = rs.random((n, 3)) < 0.05
    lo[plane] = o[plane]
    lo[rs.random((n, 3)) < 0.05] = f32(0.0)
    hi = np.maximum(hi, lo)
    ht = np.where(rs.random(n) < 0.5, f32(1e30), rs.uniform(0, 3, n).astype(f32))
    ok_rays = sign_order_rays(rcp, t0)
    assert ok_rays.sum() > n // 2
    return rcp[ok_rays], t0[ok_rays], lo[ok_rays], hi[ok_rays], ht[ok_rays]


def test_fused_record_test_passes_whenever_exact_box_passes():
    rs = np.random.default_rng(13)
    for big in (False, True):
        rcp, t0, lo, hi, ht = record_cases(rs, 200000)
        if big:  # bounds beyond the binary16 range, and the slack's 2^17 cap
            k = rs.random(lo.shape[0]) < 0.3
            lo[k] *= f32(3e5)
            hi[k] *= f32(3e5)
        eps = slab_eps(lo, hi)
        lo16, hi16 = outward(lo, True), outward(hi, False)
        tmin_e, ok_e = slab(rcp, t0, lo, hi, ht)
        tmin_w, ok_w = wide_slab(rcp, t0, lo16, hi16, ht, eps)
        assert ok_e.sum() > 10000
        assert np.all(ok_w[ok_e]), int((ok_e & ~ok_w).sum())
        assert np.all(tmin_w[ok_e] <= tmin_e[ok_e])



def slab_values(rs, n):
    """Single slab values rcp * b + t0 built to cancel (the origin next to the bound) over the
    range of reciprocals the fused test accepts, many with bounds a binary16 holds exactly."""
    mag = f32(2.0) ** rs.uniform(-100, 100, n).astype(f32)
    mag[: n // 2] = f32(2.0) ** rs.uniform(0, 14, n // 2).astype(f32)
    r = (mag * np.where(rs.random(n) < 0.5, f32(-1), f32(1))).astype(f32)
    b = rs.uniform(-8, 8, n).astype(f32)
    b[: n // 2] = np.round(b[: n // 2] * 64) / 64
    tiny = rs.random(n) < 0.05
    b[tiny] = (rs.uniform(-1, 1, tiny.sum()) * 2.0 ** -14).astype(f32)
    o = (b + rs.normal(size=n).astype(f32) * np.abs(b) * f32(1e-4)).astype(f32)
    with np.errstate(all="ignore"):
        t0 = (-r * o).astype(f32)
        ref = ((r * b).astype(f32) + t0).astype(f32)
    keep = np.abs(t0) <= f32(2.0 ** 120)
    return r[keep], b[keep], t0[keep], ref[keep], slab_eps(b, b)


def test_fused_slab_value_never_moves_inward():
    rs = np.random.default_rng(17)
    for _ in range(5):
        r, b, t0, ref, eps = slab_values(rs, 400000)
        e = slack(r, t0, eps)
        # a lower bound with r > 0 (an upper one with r < 0) is a near bound: inward means larger
        near_lo = fma32(r, outward(b, True), (t0 - e).astype(f32))
        far_lo = fma32(r, outward(b, True), (t0 + e).astype(f32))
        near_hi = fma32(r, outward(b, False), (t0 - e).astype(f32))
        far_hi = fma32(r, outward(b, False), (t0 + e).astype(f32))
        assert np.all(np.where(r > 0, near_lo <= ref, far_lo >= ref))
        assert np.all(np.where(r > 0, far_hi >= ref, near_hi <= ref))


def test_fused_slab_value_needs_the_slack():
    """Without the slack the fused value moves inward of the reference's two-rounding value on
    bounds a binary16 holds exactly (the rounding of rcp * b is then uncovered)."""
    rs = np.random.default_rng(19)
    r, b, t0, ref, eps = slab_values(rs, 400000)
    near_lo = fma32(r, outward(b, True), t0)
    assert np.any((r > 0) & (near_lo > ref))


def test_scene_records_with_the_fused_test():
    m = g.Mesh.generate(g.SCENE_SPHERES, seed=11, count=12, detail=3).build()
    b = m.arrays()["node_bounds"]
    eps = slab_eps(b[:, :3].astype(f32), b[:, 3:].astype(f32))
    rs = np.random.default_rng(23)
    idx = rs.integers(0, b.shape[0], 200000)
    # rays aimed at a point near each box (about half of them miss it)
    lo, hi = b[idx, :3].astype(f32), b[idx, 3:].astype(f32)
    aim = (lo + (hi - lo) * rs.uniform(-0.3, 1.3, lo.shape)).astype(f32)
    o = rs.uniform(-2, 2, lo.shape).astype(f32)
    d = (aim - o).astype(f32)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(f32)
    with np.errstate(all="ignore"):
        rcp = (f32(1.0) / d).astype(f32)
        t0 = (-rcp * o).astype(f32)
    ok = sign_order_rays(rcp, t0)
    rcp, t0, idx = rcp[ok], t0[ok], idx[ok]
    lo, hi = b[idx, :3].astype(f32), b[idx, 3:].astype(f32)
    ht = np.where(rs.random(idx.size) < 0.5, f32(1e30), rs.uniform(0, 2, idx.size).astype(f32))
    lo16, hi16 = outward(lo, True), outward(hi, False)
    tmin_e, ok_e = slab(rcp, t0, lo, hi, ht)
    tmin_w, ok_w = wide_slab(rcp, t0, lo16, hi16, ht, eps)
    assert ok_e.sum() > 50000
    assert np.all(ok_w[ok_e]) and np.all(tmin_w[ok_e] <= tmin_e[ok_e])
    # on a scene's boxes the slack (2^-22 of the scene's extent) admits almost nothing beyond what
    # the binary16 rounding already admits
    _, ok_r = slab(rcp, t0, lo16, hi16, ht)
    assert (ok_w & ~ok_r).sum() <= 0.002 * ok_r.sum() + 2
