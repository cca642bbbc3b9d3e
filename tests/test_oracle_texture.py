"""Textured materials: the host loader against the reference's own loader, the 512^2 x M volume
(gpurt.cpp:158-198), and the oracle's texture3D against the GL linear-filter formula in float64.

The GL driver's filtering arithmetic is not pinned by anything in the reference (parity with a
GL implementation is unpinned); what is pinned here is the GL specification's formula
(texel coordinate u*size - 1/2, floor, wrap, fractional weights), texel centres, the repeat wrap
and the quarter-slice blend of the photon lookups.
"""
import ctypes
import os

import numpy as np
import pytest

import glslppm_amd as g
import oracle


def textured_cornell(res=16, images=None, seed=5):
    """Cornell room with per-vertex texture coordinates on every triangle and images on the
    materials named in `images` ({material: (h, w, 3) uint8})."""
    m = g.Mesh.generate(g.SCENE_CORNELL).set_camera((0, 0, 1.667), (0, 0, 0), res, res, 45.0).build()
    T = m.num_triangles
    rs = np.random.default_rng(seed)
    uv = rs.uniform(-1.5, 2.5, (T, 3, 2)).astype(np.float32)
    m.set_texcoords(uv)
    if images is None:
        images = {0: rs.integers(0, 256, (7, 5, 3), dtype=np.uint8),
                  2: rs.integers(0, 256, (64, 33, 3), dtype=np.uint8)}
    for mat, img in images.items():
        m.set_texture(mat, img)
    return m.build(), uv, images


def gl_linear_f64(vol, u, v, r):
    """texture3D of vol [D][S][S][3] uint8 with GL_LINEAR / GL_REPEAT, in float64."""
    D, S = vol.shape[0], vol.shape[1]

    def axis(c, n):
        a = c * n - 0.5
        f = np.floor(a)
        i0 = int(f) % n
        return i0, (i0 + 1) % n, a - f

    i0, i1, a = axis(u, S)
    j0, j1, b = axis(v, S)
    k0, k1, c = axis(r, D)
    t = vol.astype(np.float64) / 255.0
    out = 0.0
    for ki, wk in ((k0, 1 - c), (k1, c)):
        for ji, wj in ((j0, 1 - b), (j1, b)):
            for ii, wi in ((i0, 1 - a), (i1, a)):
                out = out + wk * wj * wi * t[ki, ji, ii]
    return out


def test_volume_is_nearest_resampling_with_magenta_for_untextured():
    img = np.array([[[10, 20, 30], [40, 50, 60]], [[70, 80, 90], [100, 110, 120]]], np.uint8)  # 2 x 2
    m, _, _ = textured_cornell(images={1: img})
    a = m.arrays()
    vol = a["textures"]
    assert a["texture_size"] == 512 and vol.shape == (len(a["materials"]), 512, 512, 3)
    # quadrants of the 2x2 image: x * 2 / 512, y * 2 / 512 (gpurt.cpp:177)
    assert np.all(vol[1, :256, :256] == img[0, 0]) and np.all(vol[1, :256, 256:] == img[0, 1])
    assert np.all(vol[1, 256:, :256] == img[1, 0]) and np.all(vol[1, 256:, 256:] == img[1, 1])
    for k in range(vol.shape[0]):
        if k != 1:
            assert np.all(vol[k] == np.array([255, 0, 255], np.uint8))


def test_texcoords_only_on_textured_materials():
    m, uv, images = textured_cornell()
    a = m.arrays()
    tex = a["texcoords"]
    textured = np.isin(a["material"], list(images))
    assert np.array_equal(tex[textured], uv[textured])
    assert np.all(tex[~textured] == np.float32(1e30))


def test_texture3d_matches_the_gl_formula_in_float64():
    m, _, _ = textured_cornell()
    vol = m.arrays()["textures"]
    M = vol.shape[0]
    rs = np.random.default_rng(9)
    worst = 0.0
    for _ in range(1500):
        u, v = rs.uniform(-3, 3, 2).astype(np.float32)
        mat = int(rs.integers(0, M))
        r = np.float32(np.float32(1.0) / np.float32(M)) * np.float32(mat + rs.choice([0.5, 0.75]))
        got = oracle.texture_sample(m.scene, u, v, r).astype(np.float64)
        # the float64 reference evaluates the fp32 coordinates the oracle sees
        want = gl_linear_f64(vol, float(u), float(v), float(r))
        worst = max(worst, float(np.max(np.abs(got - want))))
    # weights come from fp32 coordinates (u*512 - 1/2 rounds to 2^-14 at |u| <= 3): a few 1e-5
    assert worst < 1e-4, worst


def test_texel_centres_are_exact():
    m, _, _ = textured_cornell()
    vol = m.arrays()["textures"]
    M = vol.shape[0]
    rs = np.random.default_rng(2)
    for _ in range(300):
        i, j = rs.integers(0, 512, 2)
        k = int(rs.integers(0, M))
        u, v = (i + 0.5) / 512.0, (j + 0.5) / 512.0
        r = np.float32(np.float32(1.0) / np.float32(M)) * np.float32(k + 0.5)
        got = oracle.texture_sample(m.scene, u, v, r)
        want = vol[k, j, i].astype(np.float32) / np.float32(255.0)
        # the material axis may carry a weight of a few ulp from 1/M's rounding
        assert np.allclose(got, want, rtol=0, atol=2e-6), (i, j, k, got, want)


def test_photon_lookup_blends_a_quarter_of_the_next_slice_with_wrap():
    """photontrace.fs:421, 564 sample at slice index + 0.75, i.e. 3/4 of slice m and 1/4 of slice
    m + 1 (wrapping to 0 for the last), as the reference's coordinate does."""
    m, _, _ = textured_cornell()
    vol = m.arrays()["textures"].astype(np.float64) / 255.0
    M = vol.shape[0]
    for k in (0, M - 1):
        r = np.float32(np.float32(1.0) / np.float32(M)) * np.float32(k + 0.75)
        got = oracle.texture_sample(m.scene, (10 + 0.5) / 512, (20 + 0.5) / 512, r)
        want = 0.75 * vol[k, 20, 10] + 0.25 * vol[(k + 1) % M, 20, 10]
        assert np.allclose(got, want, atol=1e-5), (k, got, want)


def write_textured_obj(d):
    img = (np.arange(3 * 4 * 3) * 7 % 256).astype(np.uint8).reshape(3, 4, 3)
    img[0, 0] = [32, 10, 9]  # white-space bytes: the reference's fscanf("255\n") eats them
    with open(os.path.join(d, "tex.ppm"), "wb") as f:
        f.write(b"P6\n4 3\n255\n" + img.tobytes())
    with open(os.path.join(d, "scene.mtl"), "w") as f:
        f.write("newmtl white\nKa 0 0 0\nKd 0.8 0.8 0.8\nKs 0 0 0\nNs 100\nmap_Kd tex.ppm\n")
        f.write("newmtl light\nKa 1 1 1\nKd 1 1 1\nKs 0 0 0\nNs 100\n")
    with open(os.path.join(d, "scene.obj"), "w") as f:
        f.write("mtllib scene.mtl\n")
        for p in ((-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1), (-0.2, 1, -0.2), (0.2, 1, -0.2), (0.2, 1, 0.2)):
            f.write("v %g %g %g\n" % p)
        for t in ((0, 0), (1, 0), (1, 1), (0, 1), (0.25, 0.5), (0.75, 0.5), (0.5, 0.9)):
            f.write("vt %g %g\n" % t)
        f.write("vn 0 1 0\nvn 0 -1 0\n")
        f.write("usemtl white\nf 1/1/1 3/3/1 2/2/1\nf 1/1/1 4/4/1 3/3/1\n")
        f.write("usemtl light\nf 5/5/2 6/6/2 7/7/2\n")
    return img


def test_obj_loader_reads_texcoords_and_ppm_like_the_reference(tmp_path, reference_root):
    img = write_textured_obj(str(tmp_path))
    m = g.Mesh.load_obj(str(tmp_path / "scene.obj"))
    m.set_camera((0, 0.5, 3), (0, 0.3, 0), 16, 16, 45.0).build()
    a = m.arrays()
    # white is textured: its two triangles carry their vt; the light's triangle does not
    assert np.allclose(a["texcoords"][0], [[0, 0], [1, 1], [1, 0]])
    assert np.all(a["texcoords"][2] == np.float32(1e30))
    # LoadPPM's fscanf("255\n") also skips the leading white-space bytes of the pixel data (here
    # the whole first pixel); the short read leaves zeros at the end
    data = img.tobytes().lstrip(b" \t\n\v\f\r")
    raw = np.frombuffer(data + b"\0" * (36 - len(data)), np.uint8).reshape(3, 4, 3)
    assert np.array_equal(a["textures"][0, 0, 0], raw[0, 0])
    lib = oracle.ref_library()
    if lib is None:
        pytest.skip("reference host library not built")
    n = lib.ref_load_obj(str(tmp_path).encode(), b"scene.obj", 1.0, 0.0, 0.0, 0.0)
    assert n == 3
    tc = np.zeros((3, 3, 2), np.float32)
    lib.ref_texcoords(tc.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    assert np.array_equal(tc, a["texcoords"])
    w, h = ctypes.c_int(), ctypes.c_int()
    assert lib.ref_texture_info(0, ctypes.byref(w), ctypes.byref(h)) == 1 and (w.value, h.value) == (4, 3)
    ref_img = np.zeros((3, 4, 3), np.uint8)
    lib.ref_texture(0, ref_img.ctypes.data)
    assert np.array_equal(ref_img, raw)
    assert lib.ref_texture_info(1, ctypes.byref(w), ctypes.byref(h)) == 0


def test_scene_with_texcoords_but_no_volume_is_rejected():
    m, _, _ = textured_cornell()
    s = m.scene.contents
    bad = g.Scene()
    ctypes.pointer(bad)[0] = s
    bad.textures = None
    h = ctypes.c_void_p()
    assert g.library().ppm_create(ctypes.byref(bad), ctypes.byref(g.default_params(16)), ctypes.byref(h)) == -1
