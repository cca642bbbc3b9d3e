"""Pins of the oracle's camera, emission, BRDF sampling, Russian roulette and grid set-up.

Each check is against something the reference's code does not restate: a float64 closed form
(pinhole ray, Fresnel equations, Snell's law, mirror reflection), the distribution the sampling is
meant to draw (chi-squared against the aperture disk, light area, uniform triangle points, the
cosine lobe, the Fresnel split, the survival probability of Russian roulette), exact invariants
(generator draws consumed, flux scaling, the IBL disk geometry) and a brute-force reduction (the
hit-point bounding box behind the grid scalars).  Follows eyeraytrace.fs:396-420 (camera),
photontrace.fs:345-441 (light emission), 283-341 (IBL), 443-461 (Fresnel), 487-686 (bounce),
glslppm.cpp:292-311 (grid).
"""
import math

import numpy as np
import pytest
from scipy import stats

import glslppm_amd as g
import oracle

f32 = np.float32
P_MIN = 1e-4  # fixed seeds: deterministic; a wrong distribution gives p-values far below this


def seeds(n, k=0):
    """n generator states of the reference's seed grids (xorshift128, seed floats in [0, 2^22))."""
    s, _ = oracle.seed_grid([123456789 + 7919 * k, 362436069, 521288629, 88675123], n)
    return s


def draws(rnd, n):
    r = np.array(rnd, np.float32)
    for _ in range(n):
        oracle.gpurnd(r)
    return r


def chi2_uniform(x, bins, lo=0.0, hi=1.0):
    h, _ = np.histogram(x, bins=bins, range=(lo, hi))
    return stats.chisquare(h).pvalue


def tri_area(t):
    return 0.5 * np.linalg.norm(np.cross(t[1] - t[0], t[2] - t[0]))


# ---------------------------------------------------------------- camera (eyeraytrace.fs:396-420)

def camera_scene(w=64, h=48):
    m = g.Mesh.generate(g.SCENE_CORNELL).set_camera((0.1, -0.05, 1.667), (0.02, 0.01, 0.0), w, h, 45.0).build()
    c = m.arrays()["camera"].astype(np.float64)
    return m, c[0:3], c[3:6], c[6:9], c[9:12], c[12], c[13], c[14]


def test_pinhole_camera_ray_closed_form():
    """Aperture 0: the ray leaves the camera origin along normalize(px U + py V + d W), computed in
    float64 from the same float32 pixel offsets, within 2 ulp of the unit vector (the shader's
    round trip through the focal point C + 13 d costs about one); the lens still takes its two
    generator draws; the tent filter weight is (1 - |ax|/1.5)(1 - |ay|/1.5)."""
    m, C, U, V, W, hw, hh, d = camera_scene()
    p = g.default_params(64, image_width=64, image_height=48, aperture=0.0)
    rs = np.random.default_rng(1)
    sd = seeds(400)
    worst = 0.0
    for k in range(400):
        i, j = int(rs.integers(0, 64)), int(rs.integers(0, 48))
        ax, ay = (f32(rs.uniform(-0.625, 0.625)) for _ in range(2))
        o, dr, filt, r = oracle.camera_ray(m.scene, p, i, j, ax, ay, sd[k])
        assert np.array_equal(o, C.astype(f32))
        assert np.array_equal(r, draws(sd[k], 2))
        px = float((f32(i) + f32(0.5) + ax) - f32(hw))
        py = float((f32(j) + f32(0.5) + ay) - f32(hh))
        want = px * U + py * V + d * W
        want /= np.linalg.norm(want)
        # error in units of the unit vector's ulp (2^-23): components are compared on the scale of
        # the vector, not each on its own scale
        worst = max(worst, float(np.abs(dr.astype(np.float64) - want).max() / 2.0 ** -23))
        assert filt == (f32(1.0) - abs(ax) / f32(1.5)) * (f32(1.0) - abs(ay) / f32(1.5))
    assert worst <= 2.0, worst


def test_thin_lens_samples_the_aperture_disk_uniformly_and_focuses():
    """Aperture a: ray origins are uniform on the disk of radius a about the camera origin in the
    (U, V) plane, and every ray of a pixel passes through the pixel's focal point."""
    m, C, U, V, W, hw, hh, d = camera_scene()
    a, focal = 0.1, 1.2
    p = g.default_params(64, image_width=64, image_height=48, aperture=a, focal_length=focal)
    n = 20000
    sd = seeds(n, 1)
    r2, phi = np.empty(n), np.empty(n)
    worst = 0.0
    for k in range(n):
        i, j = 17 + (k % 5), 11 + (k % 3)
        o, dr, _, _ = oracle.camera_ray(m.scene, p, i, j, 0.0, 0.0, sd[k])
        off = o.astype(np.float64) - C
        x, y = off @ U, off @ V
        assert abs(off @ W) < 1e-6
        r2[k], phi[k] = (x * x + y * y) / (a * a), math.atan2(y, x)
        px = float((f32(i) + f32(0.5)) - f32(hw))
        py = float((f32(j) + f32(0.5)) - f32(hh))
        pin = px * U + py * V + d * W
        fp = C + focal * pin / np.linalg.norm(pin)
        to = (fp - o) / np.linalg.norm(fp - o)
        worst = max(worst, float(np.linalg.norm(np.cross(to, dr.astype(np.float64)))))
    assert r2.max() <= 1.0 + 1e-5
    assert chi2_uniform(r2, 20) > P_MIN
    assert chi2_uniform(phi, 24, -math.pi, math.pi) > P_MIN
    assert worst < 2e-6, worst


# ---------------------------------------------------------------- light emission (photontrace.fs:345-441)

LIGHT_MAT = [(1.0, 1.0, 1.0), 1.7, -1, 0.0]
FLOOR_MAT = [(0.7, 0.7, 0.7), 1.7, 0, 0.0]


def lights_scene(scales=(1.0, 2.0, 3.0, 4.0)):
    """A floor and four downward-facing light triangles whose areas are in the ratio of scales."""
    tris = [np.array([[-1, -0.5, -1], [1, -0.5, 1], [1, -0.5, -1]], np.float64)]
    for k, sc in enumerate(scales):
        s = 0.05 * math.sqrt(sc)
        x0 = -0.6 + 0.35 * k
        tris.append(np.array([[x0, 0.4, 0.0], [x0 + s, 0.4, 0.0], [x0 + s, 0.4, s]], np.float64))
    pos = np.array(tris, np.float32)
    m = g.Mesh.from_arrays(pos, [FLOOR_MAT, LIGHT_MAT], [0] + [1] * len(scales))
    return m.set_camera((0, 0.2, 1.5), (0, -0.5, 0), 16, 16, 60.0).build(), pos


def test_light_pick_is_proportional_to_area():
    m, pos = lights_scene()
    a = np.array([tri_area(t.astype(np.float64)) for t in pos[1:]])
    n = 40000
    sd = seeds(n, 2)
    picks = np.array([oracle.light_sample(m.scene, sd[k])[0] for k in range(n)])
    assert set(np.unique(picks)) <= {1, 2, 3, 4}
    obs = np.array([(picks == t).sum() for t in (1, 2, 3, 4)])
    assert stats.chisquare(obs, a / a.sum() * n).pvalue > P_MIN


def test_light_points_are_uniform_on_the_triangle_and_flux_is_le_area_pi():
    m, pos = lights_scene(scales=(1.0,))
    t = pos[1].astype(np.float64)
    gn = np.cross(t[1] - t[0], t[2] - t[0])
    gn /= np.linalg.norm(gn)
    area = m.arrays()["light_area"]
    n = 40000
    sd = seeds(n, 3)
    lam = np.empty((n, 3))
    cos2, azi = np.empty(n), np.empty(n)
    ex = t[1] - t[0]
    ex /= np.linalg.norm(ex)
    ey = np.cross(gn, ex)
    for k in range(n):
        tri, flux, d, o, r = oracle.light_sample(m.scene, sd[k])
        assert tri == 1
        # flux = Le * (total light area) * 3.141592, in float32 (photontrace.fs:409-411)
        assert np.array_equal(flux, np.full(3, f32(area * f32(3.141592)), np.float32))
        # one CDF pick, two for the point, two for the direction
        assert np.array_equal(r, draws(sd[k], 5))
        q = o.astype(np.float64) - 1e-5 * gn
        A = np.stack([t[0] - t[2], t[1] - t[2]], 1)
        l01 = np.linalg.lstsq(A, q - t[2], rcond=None)[0]
        lam[k] = (l01[0], l01[1], 1 - l01.sum())
        c = float(d.astype(np.float64) @ gn)
        cos2[k] = c * c
        azi[k] = math.atan2(float(d @ ey), float(d @ ex))
        assert c > 0
    assert lam.min() > -1e-5
    # the four midpoint sub-triangles hold a quarter of the points each
    region = np.where(lam[:, 0] > 0.5, 0, np.where(lam[:, 1] > 0.5, 1, np.where(lam[:, 2] > 0.5, 2, 3)))
    assert stats.chisquare(np.bincount(region, minlength=4)).pvalue > P_MIN
    # cosine-weighted directions about the (flat) light normal: cos^2 is uniform, azimuth uniform
    assert chi2_uniform(cos2, 20) > P_MIN
    assert chi2_uniform(azi, 24, -math.pi, math.pi) > P_MIN


def test_wrong_light_pick_would_be_caught():
    """The area test has teeth: a uniform pick over the four lights fails it."""
    a = np.array([1.0, 2.0, 3.0, 4.0])
    n = 40000
    uniform = np.full(4, n / 4)
    assert stats.chisquare(uniform, a / a.sum() * n).pvalue < 1e-12


def test_ibl_emission_geometry_and_sun_split():
    """Without light triangles photons come from the environment: half along the fixed sun
    direction -normalize(1, 2, -1) with flux (1, 1, 0.7) x area, half from the sky with flux
    (0.7, 0.7, 1) x area, area = 4 pi r^2 of the bounding sphere; origins on the disk of radius r
    whose centre is r from the sphere centre, perpendicular to the direction."""
    base = g.Mesh.generate(g.SCENE_CORNELL).build().arrays()
    mats = [(rgb, eta, 0 if brdf == -1 else brdf, gl) for (rgb, eta, brdf, gl) in base["materials"]]
    m = g.Mesh.from_arrays(base["positions"], mats, base["material"], base["normals"])
    m.set_camera((0, 0, 1.667), (0, 0, 0), 16, 16, 45.0).build()
    bs = m.arrays()["bsphere"].astype(np.float64)
    area = f32(f32(12.566368) * f32(bs[3])) * f32(bs[3])
    sun = np.array([1.0, 2.0, -1.0]) / math.sqrt(6.0)
    n = 20000
    sd = seeds(n, 4)
    n_sun = 0
    for k in range(n):
        flux, d, o, _ = oracle.ibl_sample(m.scene, sd[k])
        dd = -d.astype(np.float64)
        assert abs(np.linalg.norm(dd) - 1.0) < 1e-6
        if np.allclose(dd, sun, atol=1e-6):
            n_sun += 1
            assert np.array_equal(flux, np.array([area, area, f32(0.7) * area], np.float32))
        else:
            assert np.array_equal(flux, np.array([f32(0.7) * area, f32(0.7) * area, area], np.float32))
        rel = o.astype(np.float64) - bs[:3]
        assert abs(rel @ dd - bs[3]) < 1e-5 * max(1.0, bs[3])
        assert np.linalg.norm(rel - (rel @ dd) * dd) <= bs[3] * (1 + 1e-5)
    assert stats.binomtest(n_sun, n, 0.5).pvalue > P_MIN


# ---------------------------------------------------------------- Fresnel, bounce (photontrace.fs:443-686)

def fresnel64(cos_i, n1, n2):
    s2 = (n1 / n2) ** 2 * (1 - cos_i * cos_i)
    if s2 >= 1:
        return 1.0
    cos_t = math.sqrt(1 - s2)
    rs = (n1 * cos_i - n2 * cos_t) / (n1 * cos_i + n2 * cos_t)
    rp = (n1 * cos_t - n2 * cos_i) / (n1 * cos_t + n2 * cos_i)
    return 0.5 * (rs * rs + rp * rp)


def test_fresnel_matches_the_fresnel_equations():
    rs = np.random.default_rng(5)
    worst = 0.0
    for _ in range(5000):
        c = float(rs.uniform(0.0, 1.0))
        eta = float(rs.uniform(1.05, 2.5))
        ni, ne = (1.0, eta) if rs.random() < 0.5 else (eta, 1.0)
        inc = np.array([math.sqrt(1 - c * c), c, 0.0], np.float32)
        got = float(oracle.fresnel(inc, [0.0, 1.0, 0.0], ni, ne))
        worst = max(worst, abs(got - fresnel64(float(inc[1]), ni, ne)))
    assert worst < 2e-5, worst
    # normal incidence and total internal reflection
    assert abs(float(oracle.fresnel([0, 1, 0], [0, 1, 0], 1.0, 1.5)) - 0.04) < 1e-6
    assert oracle.fresnel([0.9, math.sqrt(1 - 0.81), 0], [0, 1, 0], 1.5, 1.0) == 1.0


def bounce_scene(brdf, rgb=(0.6, 0.3, 0.45), eta=1.5):
    """One upward-facing triangle of the given material (and a light, which the bounce ignores)."""
    pos = np.array([[[-1, 0, -1], [1, 0, 1], [1, 0, -1]],
                    [[-0.1, 0.5, -0.1], [0.1, 0.5, -0.1], [0.1, 0.5, 0.1]]], np.float32)
    m = g.Mesh.from_arrays(pos, [[rgb, eta, brdf, 0.0], LIGHT_MAT], [0, 1])
    return m.set_camera((0, 0.3, 1.5), (0, 0, 0), 16, 16, 60.0).build()


def test_diffuse_bounce_cosine_lobe_and_russian_roulette():
    col = np.array([0.6, 0.3, 0.45], np.float32)
    m = bounce_scene(0, tuple(col))
    n_up = np.array([0.0, 1.0, 0.0], np.float32)
    pd = np.array([0.3, -0.9, 0.2], np.float32)
    pd /= np.linalg.norm(pd)
    flux0 = np.array([2.0, 1.5, 0.25], np.float32)
    n = 40000
    sd = seeds(n, 6)
    alive, cos2 = 0, []
    p = np.max(col)
    for k in range(n):
        cont, flux, o, d, r = oracle.photon_bounce(m.scene, 0, [0.2, 0.0, -0.1], n_up, pd, sd[k], flux0)
        assert np.array_equal(r, draws(sd[k], 3))  # two for the direction, one for the roulette
        if cont:
            alive += 1
            # surviving flux is scaled by col / max(col): unbiased in expectation
            assert np.array_equal(flux, flux0 * (col / p))
            assert d[1] >= 0 and np.array_equal(o, np.array([0.2, 1e-6, -0.1], np.float32))
            cos2.append(float(d[1]) ** 2)
    assert stats.binomtest(alive, n, float(p)).pvalue > P_MIN
    assert chi2_uniform(np.array(cos2), 20) > P_MIN


def test_mirror_bounce_reflects_and_only_rolls_the_roulette():
    col = (0.8, 0.8, 0.8)
    m = bounce_scene(1, col)
    n_up = np.array([0.0, 1.0, 0.0], np.float32)
    rs = np.random.default_rng(7)
    sd = seeds(2000, 7)
    for k in range(2000):
        pd = rs.normal(size=3)
        pd[1] = -abs(pd[1]) - 0.05
        pd = (pd / np.linalg.norm(pd)).astype(np.float32)
        cont, _, _, d, r = oracle.photon_bounce(m.scene, 0, [0.0, 0.0, 0.0], n_up, pd, sd[k], [1, 1, 1])
        assert np.array_equal(r, draws(sd[k], 1))
        if cont:
            want = pd.astype(np.float64) * np.array([1, -1, 1])
            assert np.abs(d - want).max() < 1e-6


def test_glass_bounce_fresnel_split_and_snell():
    eta = 1.5
    m = bounce_scene(2, (0.9, 0.9, 0.9), eta)
    n_up = np.array([0.0, 1.0, 0.0], np.float32)
    cos_i = 0.35
    pd = np.array([math.sqrt(1 - cos_i * cos_i), -cos_i, 0.0], np.float32)
    want_r = fresnel64(float(-pd[1]), 1.0, eta)
    n = 40000
    sd = seeds(n, 8)
    refl = refr = 0
    for k in range(n):
        cont, _, _, d, _ = oracle.photon_bounce(m.scene, 0, [0.0, 0.0, 0.0], n_up, pd, sd[k], [1, 1, 1])
        if not cont:
            continue  # a refraction keeps the glass colour, so the roulette may end it
        if d[1] > 0:
            refl += 1
            assert np.abs(d.astype(np.float64) - pd.astype(np.float64) * [1, -1, 1]).max() < 1e-6
        else:
            refr += 1
            # Snell: sin t = sin i / eta, in the plane of incidence
            assert abs(math.sqrt(1 - float(d[1]) ** 2) - float(pd[0]) / eta) < 2e-6
            assert abs(float(d[2])) < 1e-7
    # a reflection (probability R) resets the colour to white and always survives the roulette; a
    # refraction survives with max(col) = 0.9
    assert stats.binomtest(refl, n, want_r).pvalue > P_MIN
    assert stats.binomtest(refr, n, (1 - want_r) * 0.9).pvalue > P_MIN


# ---------------------------------------------------------------- grid (glslppm.cpp:292-311)

def test_grid_scalars_from_brute_force_bounds():
    """BBMin/BBMax, IR, GridScale and HashScale1 of an eye pass from a brute-force min/max over
    every hit point's stored position (misses keep (0, 0, 0))."""
    m = g.Mesh.generate(g.SCENE_SPHERES, seed=4, count=3, detail=2).set_camera((0, 0, 1.667), (0, 0, 0), 32, 32, 45.0)
    m.build()
    p = g.default_params(32, max_path_length=6)
    o = oracle.OracleRenderer(m, p)
    o.run(1)
    pos = o.readback("hit_position")[:, :3].astype(np.float32)
    mn, mx = pos.min(0), pos.max(0)
    size = f32(0.0)
    for k in range(3):
        size = f32(size + f32(mx[k] - mn[k]))
    ir = f32(f32(f32(size / f32(3.0)) / f32(32.0)) * f32(p.initial_footprint))
    gr = o.grid()
    assert f32(gr["initial_radius"]) == ir and ir > 0
    assert np.array_equal(np.array(gr["bbmin"], np.float32), (mn - ir).astype(np.float32))
    assert np.array_equal(np.array(gr["bbmax"], np.float32), (mx + ir).astype(np.float32))
    assert f32(gr["grid_scale"]) == f32(f32(0.5) / ir)
    assert f32(gr["hash_scale"]) == f32(f32(4194304.0) / f32(gr["grid_scale"]))
