"""HIP pass loop against the CPU oracle, bit for bit, through the C ABI (include/ppm.h).

Every state buffer of the pass loop is compared after a run of passes that spans several eye
passes: hit points (position, reflectance, normal, triangle, t, generator state), the
progressive statistics (tau, R, E, N), the photon state (position, flux, direction, generator
state, triangle, bounce, emitted flag), the hash (slot per photon, count and representative per
slot), the linear and tone-mapped radiance, the counters and the grid scalars.
"""
import numpy as np
import pytest

import glslppm_amd as g
from parity import ShrunkLeafBoxes, check_hash_properties, compare_all, format_report, run_pair

pytestmark = pytest.mark.gpu


def _check(mesh, params, passes, label):
    print("case %s: Q=%dx%d P=%dx%d H=%dx%d L=%d passes=%d" % (
        label, params.image_width, params.image_height, params.photon_width, params.photon_height,
        params.hash_width, params.hash_height, params.max_path_length, passes), flush=True)
    gpu, cpu = run_pair(mesh, params, passes)
    rep = compare_all(gpu, cpu)
    bad = format_report(rep)
    assert not bad, "%s: mismatching elements: %s" % (label, bad)
    return gpu, cpu


def cornell(res):
    return g.Mesh.generate(g.SCENE_CORNELL).set_camera((0, 0, 1.667), (0, 0, 0), res, res, 45.0).build()


def test_cornell_first_pass():
    _check(cornell(64), g.default_params(64, max_path_length=8), 1, "cornell-1")


def test_cornell_three_eye_passes():
    _check(cornell(64), g.default_params(64, max_path_length=8), 18, "cornell-18")


def test_spheres_glass_mirror():
    m = g.Mesh.generate(g.SCENE_SPHERES, seed=7, count=10, detail=2)
    m.set_camera((0, 0, 1.667), (0, 0, 0), 64, 64, 45.0).build()
    _check(m, g.default_params(64, max_path_length=10), 23, "spheres")


def test_leaf_boxes_that_are_not_their_triangles_boxes():
    m = g.Mesh.generate(g.SCENE_SPHERES, seed=7, count=10, detail=2)
    m.set_camera((0, 0, 1.667), (0, 0, 0), 64, 64, 45.0).build()
    _check(ShrunkLeafBoxes(m), g.default_params(64, max_path_length=10), 23, "shrunk-leaf-boxes")


def test_decoupled_sizes():
    m = g.Mesh.generate(g.SCENE_SPHERES, seed=11, count=4, detail=1)
    m.set_camera((0.1, 0.05, 1.6), (0, 0, 0), 48, 40, 55.0).build()
    p = g.default_params(64, image_width=48, image_height=40, photon_width=64, photon_height=32,
                         hash_width=128, hash_height=16, max_path_length=6)
    _check(m, p, 13, "decoupled")


def _glossy_variant(seed):
    base = g.Mesh.generate(g.SCENE_SPHERES, seed=seed, count=8, detail=2).build()
    a = base.arrays()
    mats = []
    kinds = [4, 5, 6, 3]
    for i, (rgb, eta, brdf, gl) in enumerate(a["materials"]):
        if i >= 5:  # the spheres: glossy mirror, glossy glass, glossy plastic, plastic
            brdf = kinds[i % 4]
            gl = 1.0 / ((2.0 / (1.0 - 0.6)) ** 2.71828 + 1.0)
            eta = 1.5
        mats.append((rgb, eta, brdf, gl))
    m = g.Mesh.from_arrays(a["positions"], mats, a["material"], a["normals"])
    return m.set_camera((0, 0, 1.667), (0, 0, 0), 64, 64, 45.0).build()


def test_glossy_and_plastic_brdfs():
    _check(_glossy_variant(3), g.default_params(64, max_path_length=10), 21, "glossy")


def test_image_based_light_without_light_triangles():
    base = cornell(48).arrays()
    mats = [(rgb, eta, 0 if brdf == -1 else brdf, gl) for (rgb, eta, brdf, gl) in base["materials"]]
    m = g.Mesh.from_arrays(base["positions"], mats, base["material"], base["normals"])
    m.set_camera((0, 0, 1.667), (0, 0, 0), 48, 48, 45.0).build()
    assert m.arrays()["light_triangle"].size == 0
    _check(m, g.default_params(48, hash_width=64, hash_height=32, max_path_length=5), 7, "ibl")


def test_cornell_c0_full_run():
    """256^2 eye rays and photons, <= 8 bounces, fixed initial radius 1/50 of the room, 64 passes."""
    p = g.default_params(256, max_path_length=8, fixed_radius=0.02)
    gpu, cpu = _check(cornell(256), p, 64, "c0-full")
    rad_g, rad_c = gpu.readback("radiance")[:, :3], cpu.readback("radiance")[:, :3]
    rel = np.linalg.norm(rad_g - rad_c) / np.linalg.norm(rad_c)
    assert rel <= 1e-3


def test_reference_scene_file_layout_square_512():
    """The reference's own configuration: Q = P = H = 512^2, 10 bounces, footprint radius rule."""
    m = g.Mesh.generate(g.SCENE_CORNELL).set_camera((0, 0, 1.667), (0, 0, 0), 512, 512, 45.0).build()
    _check(m, g.default_params(512), 12, "square-512")


def c2_terrain():
    """C2: 1024x1024-cell value-noise heightfield plus 8 glass icospheres (subdivision 5),
    2,260,994 triangles, small area light; 1920x1080 eye rays, 2048^2 photons, 2^24 hash slots,
    12 bounces."""
    m = g.Mesh.generate(g.SCENE_TERRAIN, seed=0x67736C72, count=8, detail=1024)
    m.set_camera((0.0, 0.55, 1.9), (0.0, -0.25, 0.0), 1920, 1080, 50.0).build()
    p = g.default_params(2048, image_width=1920, image_height=1080, hash_width=4096, hash_height=4096,
                         max_path_length=12)
    return m, p


def check_progress_properties(r, p, passes, checkpoints):
    """Runs `passes` more passes in `checkpoints` steps: at each step the hash properties hold, R
    never grows and N never shrinks at any hit point (eye passes move hit points but never reset
    the statistics), and the counters advance by exactly P per pass."""
    P = p.photon_width * p.photon_height
    c0 = r.counters()
    tr = r.readback("flux_radius").copy()
    en = r.readback("emission_count").copy()
    step = passes // checkpoints
    for k in range(checkpoints):
        r.run(step)
        check_hash_properties(r, p)
        tr1, en1 = r.readback("flux_radius").copy(), r.readback("emission_count").copy()
        assert np.all(tr1[:, 3] <= tr[:, 3]) and np.all(en1[:, 3] >= en[:, 3])
        assert np.all(np.isfinite(r.readback("radiance")))
        tr, en = tr1, en1
    c = r.counters()
    done = step * checkpoints
    assert c["frame_count"] == c0["frame_count"] + done
    assert c["photon_segments"] == c0["photon_segments"] + done * P
    assert c0["photons_emitted"] < c["photons_emitted"] <= c0["photons_emitted"] + done * P


def c1_spheres():
    """C1: the Cornell room plus 40 jittered icospheres of subdivision 4 (generator seed
    0x67736C70; 50% diffuse, 30% mirror, 20% glass), 204,836 triangles, so the BVH comes from the
    threaded build; 1024^2 eye rays, photons and hash slots, 10 bounces, the footprint radius rule.
    This is the configuration bench.py's spheres-1024 workload times."""
    m = g.Mesh.generate(g.SCENE_SPHERES, seed=0x67736C70, count=40, detail=4)
    m.set_camera((0, 0, 1.667), (0, 0, 0), 1024, 1024, 45.0).build()
    return m, g.default_params(1024, max_path_length=10)


def test_c1_spheres_full_size():
    m, p = c1_spheres()
    assert m.num_triangles == 204836
    # Bit-exact against the oracle over one eye cycle plus two passes: eye passes 0 and 10, and
    # at 2^20 pixels the second eye stage runs ahead on the side stream beside photon passes 1-9
    # (eye_overlap is on by default), so the double-buffered hit points are checked too.  Then
    # both sides run on to pass 32 (two more eye passes) and are compared again.
    gpu, cpu = _check(m, p, 12, "c1-spheres")
    assert gpu.tuning()["eye_overlap"] == 1
    gpu.run(20)
    cpu.run(20)
    rep = compare_all(gpu, cpu)
    bad = format_report(rep)
    assert not bad, "c1-spheres after 32 passes: mismatching elements: %s" % bad
    cpu.close()
    # size-independent properties through two more eye cycles on the GPU alone (passes 32-52)
    check_progress_properties(gpu, p, 20, 4)
    assert gpu.counters()["eye_samples"] == 6


def test_c2_terrain_full_size():
    m, p = c2_terrain()
    assert m.num_triangles == 2260994
    # bit-exact against the oracle over one eye cycle and one more pass: eye passes 0 and 12, the
    # second of which ran ahead beside photon passes 1-11
    gpu, cpu = _check(m, p, 13, "c2-terrain")
    cpu.close()
    # size-independent properties through the next eye cycle, GPU only (passes 13-25)
    check_progress_properties(gpu, p, 12, 3)
    c = gpu.counters()
    assert c["frame_count"] == 25 and c["eye_samples"] == 3


def test_textured_materials():
    """Textured white walls, red wall, light and mirror box (texture coordinates on every
    triangle, images of odd sizes): the eye lookup, the light-emission lookup and the photon
    bounce lookup with its quarter-slice blend, against the oracle."""
    rs = np.random.default_rng(17)
    m = cornell(64)
    T = m.num_triangles
    m.set_texcoords(rs.uniform(-1.5, 2.5, (T, 3, 2)).astype(np.float32))
    for mat, (h, w) in {0: (37, 23), 1: (8, 8), 3: (5, 9), 4: (130, 70)}.items():
        m.set_texture(mat, rs.integers(0, 256, (h, w, 3), dtype=np.uint8))
    m.build()
    a = m.arrays()
    assert a["texture_size"] == 512 and np.all(np.abs(a["texcoords"][a["material"] == 2]) == np.float32(1e30))
    _check(m, g.default_params(64, max_path_length=8), 17, "textured")


def test_thin_lens_camera():
    """Aperture > 0: the lens sample of eyeraytrace.fs:405-410 moves the ray origin."""
    m = g.Mesh.generate(g.SCENE_SPHERES, seed=13, count=6, detail=2)
    m.set_camera((0.05, 0.1, 1.5), (0, 0, 0), 48, 48, 50.0).build()
    _check(m, g.default_params(48, hash_width=64, hash_height=32, max_path_length=6, aperture=0.08,
                             focal_length=1.4), 9, "thin-lens")


def test_degenerate_sizes():
    """One pixel, one photon, one hash slot; a one-pixel image over a 4x4 hash (its hit-point box
    has zero extent, so grid_scale is inf and every gather cell is NaN); 3x1 / 1x5 strips."""
    _check(cornell(1), g.default_params(1, max_path_length=4), 9, "1x1")
    _check(cornell(1), g.default_params(1, photon_width=2, photon_height=2, hash_width=4, hash_height=4,
                                        max_path_length=4), 9, "1x1-hash4x4")
    m = g.Mesh.generate(g.SCENE_CORNELL).set_camera((0, 0, 1.667), (0, 0, 0), 3, 1, 45.0).build()
    p = g.default_params(4, image_width=3, image_height=1, photon_width=1, photon_height=5, hash_width=2,
                         hash_height=3, max_path_length=3)
    _check(m, p, 7, "strips")


def test_smallest_scenes():
    """A lone light triangle seen from below (photons leave the scene, none lands) and the same
    light over a floor."""
    light = [(1.0, 1.0, 1.0), 1.7, -1, 0.0]
    floor = [(0.7, 0.7, 0.7), 1.7, 0, 0.0]
    tri_light = np.array([[[-0.2, 0.4, -0.2], [0.2, 0.4, -0.2], [0.2, 0.4, 0.2]]], np.float32)  # faces -y
    m = g.Mesh.from_arrays(tri_light, [light], [0]).set_camera((0, -0.5, 1.0), (0, 0.4, 0), 16, 16, 60.0).build()
    _check(m, g.default_params(16, max_path_length=4), 5, "one-triangle")
    tri_floor = np.array([[[-1, -0.5, -1], [1, -0.5, 1], [1, -0.5, -1]]], np.float32)  # faces +y
    m = g.Mesh.from_arrays(np.concatenate([tri_light, tri_floor]), [light, floor], [0, 1])
    m.set_camera((0, 0.2, 1.5), (0, -0.5, 0), 16, 16, 60.0).build()
    _check(m, g.default_params(16, max_path_length=4), 9, "two-triangles")


def test_axis_parallel_rays():
    """Every eye ray exactly (0, 0, -1): the Cornell room moved to x = y = 65536 and a pinhole
    camera there with a focal distance of 1e-3, so that the focal point rounds onto the camera's
    x and y.  Such rays have infinite reciprocals on two axes, where the reference's slab test
    (eyeraytrace.fs:92-116) meets NaN; the walks must keep its outcome box by box, as they do for
    every other ray."""
    import oracle
    a = g.Mesh.generate(g.SCENE_CORNELL).build().arrays()
    off = np.array([65536.0, 65536.0, 0.0], np.float32)
    pos = (a["positions"].reshape(-1, 3, 3) + off).astype(np.float32)
    m = g.Mesh.from_arrays(pos, a["materials"], a["material"], a["normals"].reshape(-1, 3, 3))
    m.set_camera(tuple(off + np.float32([0.0, 0.0, 1.667])), tuple(off), 32, 32, 45.0).build()
    p = g.default_params(32, max_path_length=6, focal_length=1e-3, aperture=0.0)
    for (i, j) in ((0, 0), (31, 0), (16, 16), (5, 27)):
        _, d, _, _ = oracle.camera_ray(m.scene, p, i, j, 0.3, -0.4, (0.5, 0.5, 0.5, 0.5))
        assert tuple(d) == (0.0, 0.0, -1.0)
    _check(m, p, 13, "axis-parallel")


def test_duplicated_triangles_tie_by_walk_order():
    """Every non-light triangle of the Cornell room twice, the copy in another colour: every eye
    ray and every photon meets two triangles at exactly the same distance, and the reference keeps
    the one its walk tests first (strict t < closest, eyeraytrace.fs:140).  The four-wide walk with
    its deferred leaf tests must keep that same one everywhere."""
    a = g.Mesh.generate(g.SCENE_CORNELL).build().arrays()
    pos = a["positions"].reshape(-1, 3, 3).astype(np.float32)
    nrm = a["normals"].reshape(-1, 3, 3).astype(np.float32)
    mat = np.asarray(a["material"], np.int32)
    mats = list(a["materials"])
    light = np.array([m[2] == -1 for m in mats])[mat]
    extra = [((0.2, 0.3, 0.9), 1.7, 0, 0.0), ((0.9, 0.9, 0.2), 1.7, 1, 0.0)]  # diffuse blue, mirror yellow
    base = len(mats)
    dup_mat = np.where(np.array([mats[k][2] == 1 for k in mat]), base + 1, base)[~light]
    m = g.Mesh.from_arrays(np.concatenate([pos, pos[~light]]), mats + extra,
                           np.concatenate([mat, dup_mat]), np.concatenate([nrm, nrm[~light]]))
    m.set_camera((0, 0, 1.667), (0, 0, 0), 48, 48, 45.0).build()
    assert m.num_triangles == len(pos) + int((~light).sum())
    gpu, cpu = _check(m, g.default_params(48, hash_width=64, hash_height=32, max_path_length=8), 17, "duplicated")
    # both copies are hit somewhere: the tie goes to one copy or the other by walk order
    tri = gpu.readback("hit_triangle")[:, 0]
    assert (tri >= len(pos)).any() and ((tri >= 0) & (tri < len(pos))).any()
