"""Pins of the oracle's glossy perturbation and its plastic BRDFs, in photon and eye mode.

glossy_reflect (eyeraytrace.fs:370-392, photontrace.fs:258-280) perturbs a normal: the axis is
r = normalize((1 - 2g) reflect(d, n) + 2g n), the sample s = onb(v, r) has cos(s, r) = rnd2^g and
azimuth 2 pi rnd1 about r, and a sample below the surface (s . n < 0) is mirrored into the upper
hemisphere.  So (cos about the axis)^(1/g) is uniform: a Phong-like lobe of exponent 1/g - 1.  The
shaders call it with d = n, where the axis is +n or -n and the mirrored result has cos(result, n)
= rnd2^g exactly as drawn.  The plastic BRDF (photontrace.fs:639-669, eyeraytrace.fs:518-539)
reflects with the Fresnel probability R and otherwise scatters diffusely (photon mode) or stops
there as a hit point (eye mode); its glossy form (6) perturbs the normal first.

Each check is against a float64 closed form (the axis, the cosine, mirror reflection, the Fresnel
equations), a distribution (chi-squared of the lobe and the azimuth, binomial or summed-Bernoulli
tests of the Fresnel split and of Russian roulette) or an exact invariant (draws consumed, unit
length, the colour reset on reflection, flux scaling).  The generator draws come from
oracle.gpurnd, pinned in test_oracle_rng.py.  A wrong lobe (the exponent g dropped) fails the same
test: test_a_wrong_lobe_fails.
"""
import math

import numpy as np
from scipy import stats

import glslppm_amd as g
import oracle

f32 = np.float32
P_MIN = 1e-4  # fixed seeds: deterministic; a wrong distribution gives p-values far below this
G_VALUES = (f32(1.0 / ((2.0 / (1.0 - 0.6)) ** 2.71828 + 1.0)), f32(0.1), f32(0.4))  # 0.0124: specularity 0.6


def seeds(n, k=0):
    s, _ = oracle.seed_grid([123456789 + 7919 * (k + 40), 362436069, 521288629, 88675123], n)
    return s


def drawn(rnd, n):
    """The first n draws of a generator state and the state after them."""
    r = np.array(rnd, np.float32)
    out = [float(oracle.gpurnd(r)) for _ in range(n)]
    return out, r


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def reflect64(d, n):
    return d - 2.0 * (n @ d) * n


def frame(r):
    """Any orthonormal pair perpendicular to r (the test's own, not the shader's onb)."""
    a = np.array([1.0, 0.0, 0.0]) if abs(r[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = unit(np.cross(r, a))
    return e1, np.cross(r, e1)


def perturbed_normal(d, incident, cos_want):
    """The normal n' of a mirror reflection d = reflect(incident, n') = incident - 2 (n'.incident) n'
    (so n' is along +-(d - incident)), or of -d where the shader negated a reflection that pointed
    below the surface; checks that its cosine to the surface normal (+y) is the drawn cos_want.
    (A grazing ray can meet a tilted n' from behind, n'.incident > 0: hence the sign.)"""
    d, i = np.asarray(d, np.float64), np.asarray(incident, np.float64)
    cands = [unit(d - i), unit(-d - i)]
    cands = [c if c[1] >= 0 else -c for c in cands]
    err = [abs(c[1] - cos_want) for c in cands]
    assert min(err) < 3e-5, "cos(n', n) = rnd2^g: %g vs %g" % (cands[int(np.argmin(err))][1], cos_want)
    return cands[int(np.argmin(err))]


def chi2_uniform(x, bins, lo=0.0, hi=1.0):
    h, _ = np.histogram(x, bins=bins, range=(lo, hi))
    return stats.chisquare(h).pvalue


def fresnel64(cos_i, n1, n2):
    s2 = (n1 / n2) ** 2 * (1 - cos_i * cos_i)
    if s2 >= 1:
        return 1.0
    cos_t = math.sqrt(1 - s2)
    rs = (n1 * cos_i - n2 * cos_t) / (n1 * cos_i + n2 * cos_t)
    rp = (n1 * cos_t - n2 * cos_i) / (n1 * cos_t + n2 * cos_i)
    return 0.5 * (rs * rs + rp * rp)


def circular_spread(a):
    """Spread of angles about their circular mean (0 when all are equal mod 2 pi)."""
    m = math.atan2(np.sin(a).mean(), np.cos(a).mean())
    return float(np.abs(np.angle(np.exp(1j * (a - m)))).max())


def glossy_samples(d, n, gl, count, k):
    """Samples of glossy_reflect with the pre-flip sample s recovered, its cosine about the float64
    axis and its azimuth in the test's frame, next to the draws they came from."""
    d64, n64 = unit(d), unit(n)
    r64 = unit((1.0 - 2.0 * float(gl)) * reflect64(d64, n64) + 2.0 * float(gl) * n64)
    e1, e2 = frame(r64)
    sd = seeds(count, k)
    cos_r, azi, rnd1, rnd2 = np.empty(count), np.empty(count), np.empty(count), np.empty(count)
    for i in range(count):
        (u1, u2), after = drawn(sd[i], 2)
        out, r = oracle.glossy_reflect(f32(d), f32(n), gl, sd[i])
        assert np.array_equal(r, after), "glossy_reflect takes exactly two draws"
        o = out.astype(np.float64)
        assert abs(np.linalg.norm(o) - 1.0) < 2e-6
        assert o @ n64 >= -1e-7, "the result lies in the upper hemisphere of n"
        t = float(np.float64(u2) ** float(gl))
        # the sample before the flip: the result itself, or its mirror image below the surface
        cands = [o, reflect64(o, n64)]
        err = [abs(c @ r64 - t) for c in cands]
        s = cands[int(np.argmin(err))]
        assert min(err) < 2e-5, "cos(sample, axis) = rnd2^g: %g vs %g" % (s @ r64, t)
        cos_r[i], azi[i], rnd1[i], rnd2[i] = s @ r64, math.atan2(s @ e2, s @ e1), u1, u2
    return cos_r, azi, rnd1, rnd2


def test_glossy_reflect_lobe_about_the_perturbed_axis():
    """Arbitrary incident directions: the cosine about r = normalize((1-2g) reflect(d, n) + 2g n)
    is rnd2^g sample by sample, (cos)^(1/g) is uniform, and the azimuth about r is 2 pi rnd1 up to a
    fixed rotation of the frame."""
    rs = np.random.default_rng(31)
    for k, gl in enumerate(G_VALUES):
        n = unit(rs.normal(size=3))
        d = unit(rs.normal(size=3))
        if d @ n > 0:
            d = -d
        cos_r, azi, rnd1, rnd2 = glossy_samples(d, n, gl, 8000, k)
        assert chi2_uniform(np.clip(cos_r, 0, 1) ** (1.0 / float(gl)), 20) > P_MIN
        assert chi2_uniform(azi, 24, -math.pi, math.pi) > P_MIN
        ok = cos_r < 0.999  # the azimuth of a sample at the pole is not defined
        phase = 2.0 * math.pi * rnd1[ok]
        assert min(circular_spread(azi[ok] - phase), circular_spread(azi[ok] + phase)) < 2e-3


def test_glossy_reflect_as_the_shaders_call_it():
    """d = n (eyeraytrace.fs:465-521, photontrace.fs:579-641): the axis is -n for g < 1/4 and +n
    above, and either way the result's cosine to n is rnd2^g."""
    rs = np.random.default_rng(32)
    for k, gl in enumerate(G_VALUES):
        n = unit(rs.normal(size=3))
        sd = seeds(6000, 10 + k)
        cosn = np.empty(6000)
        for i in range(6000):
            (_, u2), _ = drawn(sd[i], 2)
            out, _ = oracle.glossy_reflect(f32(n), f32(n), gl, sd[i])
            cosn[i] = out.astype(np.float64) @ n
            assert abs(cosn[i] - float(np.float64(u2) ** float(gl))) < 2e-5
        assert chi2_uniform(cosn ** (1.0 / float(gl)), 20) > P_MIN


def test_a_wrong_lobe_fails():
    """The lobe test has teeth: with the exponent dropped (cos = rnd2), the same statistic
    rejects."""
    gl = G_VALUES[0]
    _, _, _, rnd2 = glossy_samples(unit([0.3, -0.8, 0.2]), [0.0, 1.0, 0.0], gl, 8000, 20)
    assert chi2_uniform(rnd2 ** (1.0 / float(gl)), 20) < 1e-12


# ---------------------------------------------------------------- photon mode (photontrace.fs:567-686)

LIGHT_MAT = [(1.0, 1.0, 1.0), 1.7, -1, 0.0]
UP = np.array([0.0, 1.0, 0.0], np.float32)


def bounce_scene(brdf, rgb, eta, gl):
    pos = np.array([[[-1, 0, -1], [1, 0, 1], [1, 0, -1]],
                    [[-0.1, 0.5, -0.1], [0.1, 0.5, -0.1], [0.1, 0.5, 0.1]]], np.float32)
    m = g.Mesh.from_arrays(pos, [[rgb, eta, brdf, float(gl)], LIGHT_MAT], [0, 1])
    return m.set_camera((0, 0.3, 1.5), (0, 0, 0), 16, 16, 60.0).build()


def test_plastic_photon_bounce_fresnel_split_lobes_and_roulette():
    col = np.array([0.7, 0.35, 0.5], np.float32)
    eta = 1.5
    m = bounce_scene(3, tuple(col), eta, 0.0)
    cos_i = 0.3
    pd = np.array([math.sqrt(1 - cos_i * cos_i), -cos_i, 0.0], np.float32)
    want_r = fresnel64(float(-pd[1]), 1.0, eta)
    flux0 = np.array([2.0, 1.5, 0.25], np.float32)
    n = 40000
    sd = seeds(n, 30)
    refl = diffuse = alive_diffuse = 0
    cos2 = []
    for k in range(n):
        cont, flux, o, d, r = oracle.photon_bounce(m.scene, 0, [0.2, 0.0, -0.1], UP, pd, sd[k], flux0)
        (u_fresnel,), _ = drawn(sd[k], 1)
        if u_fresnel < want_r - 1e-6 or (u_fresnel < want_r + 1e-6 and d[1] > 0 and abs(d[0] - pd[0]) < 1e-6):
            # reflection: the mirror direction, colour reset to white, so the roulette (its draw
            # taken all the same) always lets it through with the flux unchanged
            refl += 1
            assert np.array_equal(r, drawn(sd[k], 2)[1])
            assert cont and np.array_equal(flux, flux0)
            assert np.abs(d.astype(np.float64) - pd.astype(np.float64) * [1, -1, 1]).max() < 1e-6
        else:
            # diffuse: a cosine lobe about the normal, then Russian roulette at max(col)
            diffuse += 1
            assert np.array_equal(r, drawn(sd[k], 4)[1])
            if cont:
                alive_diffuse += 1
                assert np.array_equal(flux, flux0 * (col / np.max(col)))
                assert d[1] >= 0
                cos2.append(float(d[1]) ** 2)
    assert stats.binomtest(refl, n, want_r).pvalue > P_MIN
    assert stats.binomtest(alive_diffuse, diffuse, float(np.max(col))).pvalue > P_MIN
    assert chi2_uniform(np.array(cos2), 20) > P_MIN


def test_glossy_mirror_photon_bounce_reflects_about_the_perturbed_normal():
    gl = G_VALUES[0]
    m = bounce_scene(4, (0.8, 0.8, 0.8), 1.7, gl)
    pd = unit([0.2, -0.95, 0.1]).astype(np.float32)
    n = 20000
    sd = seeds(n, 31)
    cosn = []
    for k in range(n):
        cont, _, _, d, r = oracle.photon_bounce(m.scene, 0, [0.0, 0.0, 0.0], UP, pd, sd[k], [1, 1, 1])
        (_, u2, _), after = drawn(sd[k], 3)
        assert np.array_equal(r, after)  # two for the perturbation, one for the roulette
        if not cont:
            continue
        # d = reflect(pd, n'), or its negation where that pointed below the surface
        nprime = perturbed_normal(d, pd, float(np.float64(u2) ** float(gl)))
        cosn.append(nprime[1])
    assert chi2_uniform(np.array(cosn) ** (1.0 / float(gl)), 20) > P_MIN


def test_glossy_plastic_photon_bounce_splits_by_fresnel_at_the_perturbed_normal():
    """BRDF 6: the normal is perturbed first (the pinned glossy_reflect, two draws), then the
    plastic split uses the Fresnel term at the perturbed normal: the number of reflections matches
    the sum of those probabilities, and each reflection is the mirror direction about n'."""
    gl, eta = G_VALUES[0], 1.5
    col = np.array([0.6, 0.6, 0.3], np.float32)
    m = bounce_scene(6, tuple(col), eta, gl)
    pd = unit([0.7, -0.6, 0.2]).astype(np.float32)
    n = 30000
    sd = seeds(n, 32)
    refl, mean, var = 0, 0.0, 0.0
    for k in range(n):
        cont, flux, _, d, r = oracle.photon_bounce(m.scene, 0, [0.0, 0.0, 0.0], UP, pd, sd[k], [1, 1, 1])
        nprime, after2 = oracle.glossy_reflect(UP, UP, gl, sd[k])
        n64 = nprime.astype(np.float64)
        p_r = fresnel64(float(-(pd.astype(np.float64) @ n64)), 1.0, eta)
        mean += p_r
        var += p_r * (1.0 - p_r)
        want = reflect64(pd.astype(np.float64), n64)
        (u_fresnel,), _ = drawn(after2, 1)
        if u_fresnel < p_r - 1e-6:
            refl += 1
            assert cont and np.array_equal(flux, np.ones(3, np.float32))
            assert np.array_equal(r, drawn(sd[k], 4)[1])  # perturbation 2, Fresnel 1, roulette 1
            if want[1] > 0:
                assert np.abs(d.astype(np.float64) - want).max() < 2e-6
        elif u_fresnel > p_r + 1e-6:
            assert np.array_equal(r, drawn(sd[k], 6)[1])  # + 2 for the diffuse direction
    assert abs(refl - mean) < 4.5 * math.sqrt(var)


# ---------------------------------------------------------------- eye mode (eyeraytrace.fs:396-547)

def floor_scene(brdf, rgb, eta, gl):
    """A large floor of one material under the environment light (no light triangles), so that a
    path reflected off the floor leaves the scene."""
    pos = np.array([[[-4, 0, -4], [4, 0, 4], [4, 0, -4]], [[-4, 0, -4], [-4, 0, 4], [4, 0, 4]]], np.float32)
    m = g.Mesh.from_arrays(pos, [[rgb, eta, brdf, float(gl)]], [0, 0])
    return m.set_camera((0.0, 0.6, 1.4), (0.0, 0.0, 0.0), 16, 16, 40.0).build()


def camera_dir(m, params, i, j, ax, ay, rnd):
    _, d, filt, _ = oracle.camera_ray(m.scene, params, i, j, ax, ay, rnd)
    return d.astype(np.float64), f32(filt)


def test_plastic_eye_path_reflects_with_fresnel_or_stops():
    """BRDF 3 in an eye path: with probability R the path reflects and carries its colour on
    unchanged (here to the sky: emission (0.7, 0.7, 1) x colour, triangle -1); otherwise the
    floor becomes the hit point with colour filter x rgb."""
    rgb = np.array([0.9, 0.5, 0.2], np.float32)
    eta = 1.7
    m = floor_scene(3, tuple(rgb), eta, 0.0)
    p = g.default_params(16, max_path_length=6)
    ax, ay = f32(0.3), f32(-0.2)
    n = 20000
    sd = seeds(n, 40)
    cam, filt = camera_dir(m, p, 8, 5, ax, ay, sd[0])
    want_r = fresnel64(float(-cam[1]), 1.0, eta)
    refl = 0
    for k in range(n):
        e, r = oracle.eye_path(m.scene, p, 8, 5, ax, ay, sd[k])
        (_, _, u), after = drawn(sd[k], 3)  # two for the lens, one for the Fresnel choice
        assert np.array_equal(r, after)
        if e["tri"] < 0:
            refl += 1
            assert u < want_r + 1e-6
            assert np.array_equal(e["col"], np.full(3, filt, np.float32))
            assert np.array_equal(e["emission"], np.array([0.7, 0.7, 1.0], np.float32) * e["col"])
            assert np.abs(e["dir"].astype(np.float64) - reflect64(cam, np.array([0.0, 1.0, 0.0]))).max() < 1e-6
        else:
            assert u > want_r - 1e-6
            assert np.array_equal(e["col"], np.full(3, filt, np.float32) * rgb)
            assert abs(float(e["pos"][1])) < 1e-6 and np.array_equal(e["nrm"], UP)
    assert stats.binomtest(refl, n, want_r).pvalue > P_MIN


def test_glossy_mirror_eye_path_reflects_about_the_perturbed_normal():
    gl = G_VALUES[0]
    rgb = np.array([0.8, 0.8, 0.8], np.float32)
    m = floor_scene(4, tuple(rgb), 1.7, gl)
    p = g.default_params(16, max_path_length=6)
    ax, ay = f32(-0.1), f32(0.25)
    n = 15000
    sd = seeds(n, 41)
    cam, filt = camera_dir(m, p, 7, 9, ax, ay, sd[0])
    cosn = []
    for k in range(n):
        e, r = oracle.eye_path(m.scene, p, 7, 9, ax, ay, sd[k])
        (_, _, _, u2), after = drawn(sd[k], 4)  # lens 2, perturbation 2
        assert np.array_equal(r, after)
        assert e["tri"] < 0 and np.array_equal(e["col"], np.full(3, filt, np.float32) * rgb)
        # d = reflect(cam, n'), or its negation where that pointed below the floor
        nprime = perturbed_normal(e["dir"], cam, float(np.float64(u2) ** float(gl)))
        cosn.append(nprime[1])
    assert chi2_uniform(np.array(cosn) ** (1.0 / float(gl)), 20) > P_MIN


def test_glossy_plastic_eye_path_splits_by_fresnel_at_the_perturbed_normal():
    gl, eta = G_VALUES[0], 1.7
    rgb = np.array([0.3, 0.6, 0.9], np.float32)
    m = floor_scene(6, tuple(rgb), eta, gl)
    p = g.default_params(16, max_path_length=6)
    ax, ay = f32(0.0), f32(0.0)
    n = 20000
    sd = seeds(n, 42)
    cam, filt = camera_dir(m, p, 8, 3, ax, ay, sd[0])
    refl, mean, var = 0, 0.0, 0.0
    for k in range(n):
        e, r = oracle.eye_path(m.scene, p, 8, 3, ax, ay, sd[k])
        _, lens = drawn(sd[k], 2)
        nprime, after = oracle.glossy_reflect(UP, UP, gl, lens)
        n64 = nprime.astype(np.float64)
        p_r = fresnel64(float(-(cam @ n64)), 1.0, eta)
        mean += p_r
        var += p_r * (1.0 - p_r)
        (u,), after = drawn(after, 1)
        assert np.array_equal(r, after)  # lens 2, perturbation 2, Fresnel 1
        if e["tri"] < 0:
            refl += 1
            assert u < p_r + 1e-6 and np.array_equal(e["col"], np.full(3, filt, np.float32))
            want = reflect64(cam, n64)
            if want[1] > 0:
                assert np.abs(e["dir"].astype(np.float64) - want).max() < 2e-6
        else:
            assert u > p_r - 1e-6 and np.array_equal(e["col"], np.full(3, filt, np.float32) * rgb)
            assert np.array_equal(e["nrm"], nprime)  # the hit point keeps the perturbed normal
    assert abs(refl - mean) < 4.5 * math.sqrt(var)
