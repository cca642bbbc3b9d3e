"""Physics sanity of the whole pass loop on the GPU: the converged radiance of an analytic scene.

A diffuse floor (albedo rho) under one small square area light of radiance Le facing down, no
other geometry: every photon that lands comes straight from the light, so each hit point's
irradiance is E(x) = Le * integral over the light of cos(theta_l) cos(theta_x) / r^2 dA, and the
floor's radiance is rho / pi * E(x).  The reference weights each eye pass's hit-point
reflectance by the tent filter of that pass's anti-aliasing offset, (1 - |ax|/1.5)(1 - |ay|/1.5)
(eyeraytrace.fs:412), without normalising it, so its estimate converges to fbar * rho / pi * E(x)
with fbar the mean filter weight of the eye passes that ran (0.627 in expectation for offsets
uniform in +-0.625).  The estimate over a patch of pixels must agree within the tolerance below.
(The emitted power per photon, Le * A * pi, is pinned in test_oracle_passes.py.)
"""
import numpy as np
import pytest

import glslppm_amd as g
import oracle

pytestmark = pytest.mark.gpu

RHO = 0.5
H = 1.0      # light height above the floor
SIDE = 0.1   # light side


def analytic_scene(res):
    s = SIDE / 2
    light = np.array([[[-s, H, -s], [s, H, -s], [s, H, s]],
                      [[-s, H, -s], [s, H, s], [-s, H, s]]], np.float32)  # faces -y
    floor = np.array([[[-3, 0, -3], [3, 0, 3], [3, 0, -3]],
                      [[-3, 0, -3], [-3, 0, 3], [3, 0, 3]]], np.float32)  # faces +y
    for t in light:
        assert np.cross(t[1] - t[0], t[2] - t[0])[1] < 0
    for t in floor:
        assert np.cross(t[1] - t[0], t[2] - t[0])[1] > 0
    mats = [[(RHO, RHO, RHO), 1.7, 0, 0.0], [(1.0, 1.0, 1.0), 1.7, -1, 0.0]]
    m = g.Mesh.from_arrays(np.concatenate([floor, light]), mats, [0, 0, 1, 1])
    # oblique camera on the floor around the origin, clear of the light
    return m.set_camera((0.9, 1.2, 0.9), (0.0, 0.0, 0.0), res, res, 15.0).build()


def irradiance(x, n=64):
    """E at floor points x [k, 3] from the light square, midpoint rule over n x n cells (Le = 1)."""
    s = SIDE / 2
    u = (np.arange(n) + 0.5) / n * SIDE - s
    ly, lx = np.meshgrid(u, u, indexing="ij")
    L = np.stack([lx.ravel(), np.full(lx.size, H), ly.ravel()], 1)  # light cell centres
    dA = (SIDE / n) ** 2
    d = L[None, :, :] - x[:, None, :]
    r2 = (d * d).sum(-1)
    cos_x = d[..., 1] / np.sqrt(r2)   # floor normal +y
    cos_l = d[..., 1] / np.sqrt(r2)   # light normal -y, seen from below
    return (cos_x * cos_l / r2).sum(-1) * dA


def mean_filter_weight(p, eye_passes):
    """The tent filter weights of the eye passes' AA offsets: the XORShift stream after the eye and
    photon seed grids gives two draws per eye pass (glslppm.cpp:113-140, 275-278)."""
    st = list(p.xorshift)
    _, st = oracle.seed_grid(st, p.image_width * p.image_height)
    _, st = oracle.seed_grid(st, p.photon_width * p.photon_height)
    w = []
    for _ in range(eye_passes):
        ax, st = oracle.xorshift_frand(st)
        ay, st = oracle.xorshift_frand(st)
        ax = (ax - np.float32(0.5)) * np.float32(1.25)
        ay = (ay - np.float32(0.5)) * np.float32(1.25)
        w.append(float((np.float32(1.0) - abs(ax) / np.float32(1.5)) * (np.float32(1.0) - abs(ay) / np.float32(1.5))))
    return float(np.mean(w))


def test_converged_radiance_of_a_floor_under_a_small_light():
    res, passes, L = 64, 800, 4
    m = analytic_scene(res)
    p = g.default_params(512, image_width=res, image_height=res, hash_width=1024, hash_height=1024,
                         max_path_length=L, fixed_radius=0.03)
    r = g.Renderer(m, p)
    r.run(passes)
    tri = r.readback("hit_triangle")[:, 0]
    pos = r.readback("hit_position")[:, :3].astype(np.float64)
    rad = r.readback("radiance")[:, :3].astype(np.float64)
    patch = np.zeros((res, res), bool)
    patch[res // 2 - 8:res // 2 + 8, res // 2 - 8:res // 2 + 8] = True
    sel = patch.ravel() & (tri >= 0) & (tri <= 1)  # floor pixels of the central patch
    assert sel.sum() > 200
    fbar = mean_filter_weight(p, passes // L)
    want = fbar * RHO / np.pi * irradiance(pos[sel])
    got = rad[sel]
    assert np.allclose(got[:, 0], got[:, 1]) and np.allclose(got[:, 0], got[:, 2])  # grey
    ratio = got[:, 0].sum() / want.sum()
    print("patch radiance / analytic = %.4f over %d pixels (mean filter weight %.4f)" % (ratio, sel.sum(), fbar),
          flush=True)
    assert abs(ratio - 1.0) < 0.03
