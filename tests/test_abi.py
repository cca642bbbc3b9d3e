"""The C ABI library: exports, argument checking and independence from the oracle (no GPU needed)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import glslppm_amd as g

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ppm.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(ppm_[a-z_0-9]+)\s*\(", text)
    return sorted(set(names))


def exported_symbols(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_every_declared_function_is_exported():
    names = declared_functions()
    assert len(names) >= 20
    syms = exported_symbols(g.LIB_PATH)
    missing = [n for n in names if n not in syms]
    assert not missing, missing
    lib = g.library()
    for n in names:
        assert getattr(lib, n) is not None


def test_product_library_does_not_link_the_oracle():
    syms = exported_symbols(g.LIB_PATH)
    assert not any("oracle" in s for s in syms)
    deps = subprocess.run(["readelf", "-d", g.LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert "oracle" not in deps and "libref_host" not in deps


def _scene():
    return g.Mesh.generate(g.SCENE_CORNELL).set_camera((0, 0, 1.667), (0, 0, 0), 32, 32, 45.0).build()


def _create(mesh, params):
    h = ctypes.c_void_p()
    rc = g.library().ppm_create(mesh.scene, ctypes.byref(params), ctypes.byref(h))
    if rc == 0:
        g.library().ppm_destroy(h)
    return rc


@pytest.mark.parametrize("field,value", [
    ("image_width", 0), ("photon_height", -1), ("hash_width", 48), ("hash_width", 8192),
    ("max_path_length", 0), ("alpha", 0.0), ("alpha", 1.5), ("initial_footprint", 0.0), ("fixed_radius", -1.0),
])
def test_create_rejects_bad_parameters(field, value):
    m = _scene()
    p = g.default_params(32)
    if field == "hash_width" and value == 8192:
        p.hash_height = 4096  # H = 2^25 > 2^24
    setattr(p, field, value)
    assert _create(m, p) == -1
    assert g.library().ppm_last_error(None)


def test_create_rejects_bad_scene():
    m = _scene()
    p = g.default_params(32)
    s = m.scene.contents
    bad = g.Scene()
    ctypes.pointer(bad)[0] = s
    bad.num_nodes = s.num_nodes - 1
    h = ctypes.c_void_p()
    assert g.library().ppm_create(ctypes.byref(bad), ctypes.byref(p), ctypes.byref(h)) == -1
    assert g.library().ppm_create(None, ctypes.byref(p), ctypes.byref(h)) == -1
    assert g.library().ppm_create(m.scene, None, ctypes.byref(h)) == -1


def test_null_context_calls_fail_cleanly():
    lib = g.library()
    assert lib.ppm_pass(None, 1) == -1
    assert lib.ppm_sync(None) == -1
    assert lib.ppm_buffer_bytes(None, 0) == 0
    buf = np.zeros(16, np.uint8)
    assert lib.ppm_readback(None, 0, buf.ctypes.data, 16) == -1
    assert lib.ppm_last_pass_ms(None) == -1.0
    lib.ppm_destroy(None)


def test_create_without_gpu_reports_no_device():
    count = ctypes.c_int(0)
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        rc = hip.hipGetDeviceCount(ctypes.byref(count))
    except OSError:
        rc = 1
    if rc == 0 and count.value > 0:
        pytest.skip("a GPU is present")
    assert _create(_scene(), g.default_params(32)) == -5


def test_default_params_are_the_reference_constants():
    p = g.default_params(512)
    assert (p.image_width, p.photon_width, p.hash_width) == (512, 512, 512)
    assert p.max_path_length == 10 and p.alpha == np.float32(0.7) and p.initial_footprint == 2.5
    assert p.focal_length == 13.0 and p.aperture == 0.0
    assert list(p.xorshift) == [123456789, 362436069, 521288629, 88675123]


def test_unknown_flag_is_rejected():
    p = g.default_params(32)
    p.flags = 4
    assert _create(_scene(), p) == -1


def test_null_batch_calls_fail_cleanly():
    lib = g.library()
    assert lib.ppm_batch_pass(None, 1) == -1
    assert lib.ppm_batch_sync(None) == -1
    assert lib.ppm_batch_remove(None, 0) == -1
    assert lib.ppm_batch_set_passes(None, 0, 1) == -1
    assert lib.ppm_batch_buffer_bytes(None, 0, 0) == 0
    assert lib.ppm_batch_num_jobs(None) == 0
    assert lib.ppm_batch_device_bytes(None) == 0
    assert lib.ppm_batch_last_pass_ms(None) == -1.0
    buf = np.zeros(16, np.uint8)
    assert lib.ppm_batch_readback(None, 0, 0, buf.ctypes.data, 16) == -1
    assert lib.ppm_batch_create(None, None) == -1
    lib.ppm_batch_destroy(None)


def test_batch_create_without_gpu_reports_no_device():
    count = ctypes.c_int(0)
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        rc = hip.hipGetDeviceCount(ctypes.byref(count))
    except OSError:
        rc = 1
    if rc == 0 and count.value > 0:
        pytest.skip("a GPU is present")
    cfg = g.BatchConfig()
    g.library().ppm_batch_config_default(ctypes.byref(cfg))
    assert cfg.passes_per_graph == 0  # plain launches, the eye stages run ahead
    h = ctypes.c_void_p()
    assert g.library().ppm_batch_create(ctypes.byref(cfg), ctypes.byref(h)) == -5
    assert g.library().ppm_batch_last_error(None)


def test_scene_footprint_counts_every_state_array():
    """Per-job HBM: 252 B per pixel (the eye pass's outputs twice: 2 x (5 x 16 + 4 + 8); the
    progressive statistics 32; RGBA8 4; resolve targets 32), 124 B per photon slot (position, flux,
    direction twice, one half per pass parity, 6 x 16; generator 16; state 8; slot 4), 24 B per
    hash slot (three tables of count and representative, used in turn), plus the scene and small
    tables, each array rounded up to the pool's block size (at most 1/8 above)."""
    m = _scene()
    small = g.scene_footprint(m, g.default_params(32))
    p = g.default_params(1024)
    big = g.scene_footprint(m, p)
    Q = P = H = 1024 * 1024
    exact = 252 * Q + 124 * P + 24 * H
    assert exact <= big <= exact * 1.125 + 64 * 65536
    assert small < big


def test_tuning_calls_check_their_arguments():
    lib = g.library()
    t = g.Tuning()
    assert lib.ppm_get_tuning(None, ctypes.byref(t)) == -1
    assert lib.ppm_set_tuning(None, ctypes.byref(t)) == -1
    assert lib.ppm_autotune(None, 1, ctypes.byref(t)) == -1
    assert lib.ppm_batch_get_tuning(None, ctypes.byref(t)) == -1
    assert lib.ppm_batch_set_tuning(None, ctypes.byref(t)) == -1
    assert lib.ppm_batch_autotune(None, 1, ctypes.byref(t)) == -1


def test_null_server_calls_fail_cleanly():
    lib = g.library()
    sid = ctypes.c_int64()
    assert lib.ppm_server_add_scene(None, None, ctypes.byref(sid)) == -1
    assert lib.ppm_server_remove_scene(None, 1) == -1
    assert lib.ppm_server_submit(None, 1, None, None, 1, ctypes.byref(sid)) == -1
    assert lib.ppm_server_wait(None, 1, 0.0) == -1
    info = g.JobInfo()
    assert lib.ppm_server_info(None, 1, ctypes.byref(info)) == -1
    buf = np.zeros(16, np.uint8)
    assert lib.ppm_server_result(None, 1, 0, buf.ctypes.data, 16) == -1
    assert lib.ppm_server_release(None, 1) == -1
    st = g.ServerStats()
    assert lib.ppm_server_stats_get(None, ctypes.byref(st)) == -1
    assert lib.ppm_server_now(None) == -1.0
    lib.ppm_server_destroy(None)


def test_server_config_defaults_and_no_device():
    cfg = g.ServerConfig()
    g.library().ppm_server_config_default(ctypes.byref(cfg))
    assert cfg.hbm_budget == 200 << 30 and cfg.max_jobs == 256 and cfg.passes_per_step == 8
    assert cfg.passes_per_graph == 0
    bad = g.ServerConfig()
    g.library().ppm_server_config_default(ctypes.byref(bad))
    bad.passes_per_step = 0
    h = ctypes.c_void_p()
    assert g.library().ppm_server_create(ctypes.byref(bad), ctypes.byref(h)) == -1
    count = ctypes.c_int(0)
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        rc = hip.hipGetDeviceCount(ctypes.byref(count))
    except OSError:
        rc = 1
    if rc == 0 and count.value > 0:
        pytest.skip("a GPU is present")
    assert g.library().ppm_server_create(ctypes.byref(cfg), ctypes.byref(h)) == -5


def test_scene_with_a_triangle_in_two_leaves_or_none_is_rejected():
    """Every triangle must sit in exactly one leaf (its exact box is packed with the triangle)."""
    m = _scene()
    nt = m.arrays()["node_triangle"].astype(np.int32)
    leaves = np.nonzero(nt >= 0)[0]
    dup = nt.copy()
    dup[leaves[1]] = dup[leaves[0]]  # triangle in two leaves, another in none
    none = nt.copy()
    none[leaves[0]] = -1             # a triangle in no leaf
    for arr, msg in ((dup, b"triangle in two leaves"), (none, b"triangle in no leaf")):
        s = g.Scene()
        ctypes.pointer(s)[0] = m.scene.contents
        s.node_triangle = arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        h = ctypes.c_void_p()
        assert g.library().ppm_create(ctypes.byref(s), ctypes.byref(g.default_params(16)), ctypes.byref(h)) == -1
        assert g.library().ppm_last_error(None) == msg
