"""ppm_pass_static, the pass loop that allocates nothing, against ppm_pass, bit for bit.

ppm_pass_static runs the same passes as ppm_pass with the passes' partial results (the eye
kernel's per-block min/max of hit positions) in the library's static device scratch,
which keeps the last call's contents.  Two calls in a row on different contexts, the second
smaller, must each give ppm_pass's results; so must a context that alternates the two entry
points.  A context whose partial results do not fit is refused before anything is queued.
"""
import ctypes

import pytest

import glslppm_amd as g
from parity import STATE_BUFFERS, compare_all, format_report

pytestmark = pytest.mark.gpu


def spheres(seed, w, h):
    m = g.Mesh.generate(g.SCENE_SPHERES, seed=seed, count=5, detail=2)
    return m.set_camera((0.1, 0.05, 1.6), (0, 0, 0), w, h, 55.0).build()


def check(a, b, label):
    rep = compare_all(b, a, STATE_BUFFERS)
    bad = format_report(rep)
    assert not bad, "%s: ppm_pass_static differs from ppm_pass: %s" % (label, bad)


def test_static_passes_equal_ppm_pass_on_two_calls_the_second_smaller():
    for label, (seed, w, h, pw, ph, L, n) in (("larger", (3, 96, 80, 128, 96, 6, 14)),
                                              ("smaller", (5, 40, 24, 32, 16, 4, 9))):
        print("case %s: %dx%d image, %dx%d photons, L=%d, %d passes" % (label, w, h, pw, ph, L, n), flush=True)
        m = spheres(seed, w, h)
        p = g.default_params(64, image_width=w, image_height=h, photon_width=pw, photon_height=ph, max_path_length=L)
        a = g.Renderer(m, p)
        b = g.Renderer(m, p)
        a.run(n)
        b.run_static(n)
        b.sync()
        check(a, b, label)
        a.close()
        b.close()


def test_static_and_plain_passes_alternate_on_one_context():
    """With the eye stage run ahead on the second stream (eye_overlap on), calls of the two entry
    points alternate across eye passes; graph replay stays on for the ppm_pass calls."""
    m = spheres(7, 64, 48)
    p = g.default_params(64, image_height=48, max_path_length=3, flags=g.FLAG_GRAPH)
    a = g.Renderer(m, p)
    b = g.Renderer(m, p)
    b.set_tuning(eye_overlap=1, passes_per_graph=1)
    for k, n in enumerate((1, 2, 4, 3, 5)):
        a.run(n)
        (b.run_static if k % 2 == 0 else b.run)(n)
        b.sync()
        check(a, b, "alternating, step %d" % k)


def test_contexts_too_large_for_the_scratch_are_refused():
    """A 6144x2048 image has 49,152 tiles of 16x16 pixels, whose min/max partials (24 bytes each)
    exceed the 1 MiB scratch."""
    m = spheres(9, 6144, 2048)
    p = g.default_params(64, image_width=6144, image_height=2048, max_path_length=4)
    r = g.Renderer(m, p)
    with pytest.raises(g.PPMError, match="static scratch"):
        r.run_static(1)
    assert r.counters()["frame_count"] == 0
    r.close()
    lib = g.library()
    assert lib.ppm_pass_static(ctypes.c_void_p(), 1) == -1
