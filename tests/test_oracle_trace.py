"""The oracle's BVH walk (eyeraytrace.fs:67-254, photontrace.fs:123-218) against brute force.

For every ray the closest hit over all triangles is computed with the textbook Moller-Trumbore
test in float32 numpy arithmetic, in the shader's operation order; the threaded walk must return
the same triangle and the same t.  Rays whose two best t are equal (edge and vertex ties) are
counted and allowed to differ in triangle only.  Eye mode culls back faces (det > 1e-10); photon
mode skips light triangles.
"""
import numpy as np

import glslppm_amd as g
import oracle

f32 = np.float32


def brute_force(pos, brdf_of_tri, org, d, cull, photon):
    v0, v1, v2 = pos[:, 0], pos[:, 1], pos[:, 2]
    e0 = (v1 - v0).astype(f32)
    e1 = (v2 - v0).astype(f32)
    d = d.astype(f32)
    org = org.astype(f32)
    pv = np.stack([d[1] * e1[:, 2] - d[2] * e1[:, 1], d[2] * e1[:, 0] - d[0] * e1[:, 2],
                   d[0] * e1[:, 1] - d[1] * e1[:, 0]], 1).astype(f32)
    det = (e0[:, 0] * pv[:, 0] + e0[:, 1] * pv[:, 1] + e0[:, 2] * pv[:, 2]).astype(f32)
    tv = (org[None, :] - v0).astype(f32)
    qv = np.stack([tv[:, 1] * e0[:, 2] - tv[:, 2] * e0[:, 1], tv[:, 2] * e0[:, 0] - tv[:, 0] * e0[:, 2],
                   tv[:, 0] * e0[:, 1] - tv[:, 1] * e0[:, 0]], 1).astype(f32)
    with np.errstate(divide="ignore", invalid="ignore"):
        u = ((tv[:, 0] * pv[:, 0] + tv[:, 1] * pv[:, 1] + tv[:, 2] * pv[:, 2]).astype(f32) / det).astype(f32)
        v = ((d[0] * qv[:, 0] + d[1] * qv[:, 1] + d[2] * qv[:, 2]).astype(f32) / det).astype(f32)
        t = ((e1[:, 0] * qv[:, 0] + e1[:, 1] * qv[:, 1] + e1[:, 2] * qv[:, 2]).astype(f32) / det).astype(f32)
    w = (f32(1.0) - u - v).astype(f32)
    ok = (u >= 0) & (v >= 0) & (t >= 0) & (w >= 0)
    if cull:
        ok &= det > f32(1e-10)
    if photon:
        ok &= brdf_of_tri != -1
    limit = f32(1e30) if photon else f32(1e20)
    ok &= t < limit
    if not ok.any():
        return -1, limit, 0
    tt = np.where(ok, t, np.inf)
    best = int(np.argmin(tt))
    ties = int(np.sum(tt == tt[best]))
    return best, tt[best], ties


def _scenes():
    yield g.Mesh.generate(g.SCENE_CORNELL).build()
    yield g.Mesh.generate(g.SCENE_SPHERES, seed=21, count=4, detail=2).build()


def test_walk_equals_brute_force_closest_hit():
    rng = np.random.default_rng(7)
    checked = ties_seen = 0
    for mesh in _scenes():
        a = mesh.arrays()
        pos = a["positions"]
        brdf_of_tri = np.array([a["materials"][m][2] for m in a["material"]])
        for k in range(1500):
            org = rng.uniform(-0.45, 0.45, 3).astype(f32)
            if k % 3 == 0:
                org = np.array([0, 0, 1.667], f32)  # eye position outside the open room
            d = rng.standard_normal(3).astype(f32)
            d = (d / np.float32(np.sqrt(np.sum(d * d)))).astype(f32)
            for cull, photon in ((False, False), (True, False), (False, True)):
                tri, t, _ = oracle.trace(mesh.scene, org, d, cull, photon)
                btri, bt, ties = brute_force(pos, brdf_of_tri, org, d, cull, photon)
                if ties > 1:
                    ties_seen += 1
                    assert t == bt
                    continue
                assert tri == btri, (org, d, cull, photon, tri, btri)
                assert t == bt
                checked += 1
    assert checked > 8000
    # exact ties are coplanar overlaps (box bottoms lie on the floor) and shared edges
    assert ties_seen < checked * 1e-2


def test_axis_parallel_rays_miss_like_the_slab_test():
    # A zero direction component makes 1/d infinite: for that axis the slab interval is
    # (inf * lo - inf, inf * hi - inf) = (-inf, NaN) for an origin inside the slab, and with
    # either GLSL's min/max or NaN-rejecting ones tmax becomes -inf, so the root box is missed.
    # The reference's walk (eyeraytrace.fs:98-105) therefore loses exactly axis-parallel rays;
    # the oracle keeps that behaviour.  A tiny tilt restores the brute-force answer.
    mesh = g.Mesh.generate(g.SCENE_CORNELL).build()
    a = mesh.arrays()
    brdf_of_tri = np.array([a["materials"][m][2] for m in a["material"]])
    for d in ([1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]):
        for org in ([0.01, 0.02, 0.03], [-0.3, 0.2, 0.1]):
            o, dd = np.array(org, f32), np.array(d, f32)
            tri, t, _ = oracle.trace(mesh.scene, o, dd, False, True)
            assert tri == -1 and t == f32(1e30)
            tilted = (dd + f32(1e-3) * np.array([1, 1, 1], f32)).astype(f32)
            tilted = (tilted / np.float32(np.sqrt(np.sum(tilted * tilted)))).astype(f32)
            tri, t, _ = oracle.trace(mesh.scene, o, tilted, False, True)
            btri, bt, ties = brute_force(a["positions"], brdf_of_tri, o, tilted, False, True)
            assert t == bt and (ties > 1 or tri == btri)
