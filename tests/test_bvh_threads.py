"""The threaded BVH build against the reference's TBVH::Build (bvh.cpp:244-317) at the sizes that
use it.

Above 1024 triangles per range the product builds disjoint subtrees on a worker pool, above
262,144 triangles it also splits each axis sort across threads, and above 65,536 nodes it threads
the six link orders.  Exact ties of the walks at C1 and C2 are decided by the tree, so it must be
the reference's, bit for bit, whatever the thread count.  tests/golden/reference_bvh_digests.json
holds SHA-256 digests of the reference's own trees (written by make_bvh_digests.py from the
reference's bvh.cpp compiled into oracle/_ref); where that build is present here, the smallest
case is also compared array by array, live.
"""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

import glslppm_amd as g
import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_bvh_digests.json")


def _cases():
    with open(GOLDEN) as f:
        return {c["name"]: c for c in json.load(f)}


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _build(case, threads, monkeypatch):
    monkeypatch.setenv("PPM_BUILD_THREADS", str(threads))
    m = g.Mesh.generate(case["kind"], seed=case["seed"], count=case["count"], detail=case["detail"]).build()
    assert m.num_triangles == case["triangles"]
    return m


def _check(case, a):
    assert a["node_bounds"].shape[0] == case["nodes"]
    assert _digest(a["node_bounds"]) == case["bounds"], "node bounds differ from the reference's tree"
    assert _digest(a["node_triangle"]) == case["tri"], "leaf triangles differ from the reference's tree"
    assert _digest(a["node_links"]) == case["links"], "hit/miss links differ from the reference's tree"


@pytest.mark.parametrize("name", ["spheres-20k", "spheres-c1", "spheres-307k"])
@pytest.mark.parametrize("threads", [1, 8, 16])
def test_threaded_build_equals_reference(name, threads, monkeypatch):
    case = _cases()[name]
    m = _build(case, threads, monkeypatch)
    _check(case, m.arrays())
    m.close()


def test_c2_terrain_tree_equals_reference(monkeypatch):
    """2,260,994 triangles: every threaded path at once (the sorts above 262,144 triangles run on
    several threads at the top levels)."""
    case = _cases()["terrain-c2"]
    m = _build(case, 16, monkeypatch)
    _check(case, m.arrays())
    m.close()


def test_threaded_build_live_against_reference_build(monkeypatch):
    """The 20k-triangle case against the reference's build run here, array by array, so that a
    mismatch names the array and the first differing node."""
    lib = oracle.ref_library()
    if lib is None:
        pytest.skip("reference host build (oracle/_ref) not present")
    case = _cases()["spheres-20k"]
    m = _build(case, 8, monkeypatch)
    a = m.arrays()
    T = a["positions"].shape[0]
    pos = np.ascontiguousarray(a["positions"].reshape(T, 3, 3), np.float32)
    nrm = np.ascontiguousarray(a["normals"].reshape(T, 3, 3), np.float32)
    mi = np.ascontiguousarray(a["material"], np.int32)
    brdf = np.array([mm[2] for mm in a["materials"]], np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    lib.ref_mesh_from_arrays(T, oracle._fptr(pos), oracle._fptr(nrm), mi.ctypes.data_as(ip), len(brdf),
                             brdf.ctypes.data_as(ip))
    rb, rt, rl = oracle.ref_bvh(lib)
    for name, ours, ref in (("node_bounds", a["node_bounds"], rb), ("node_triangle", a["node_triangle"], rt),
                            ("node_links", a["node_links"], rl)):
        diff = np.nonzero(np.any((ours != ref).reshape(ours.shape[0], -1), axis=1))[0] if name != "node_links" else \
            np.nonzero(np.any(ours != ref, axis=(0, 2)))[0]
        assert diff.size == 0, "%s differs first at node %d" % (name, diff[0])
    m.close()


def test_a_changed_tree_fails_the_digest(monkeypatch):
    """The digest check is not vacuous: swapping two leaves' triangles changes it."""
    case = _cases()["spheres-20k"]
    m = _build(case, 8, monkeypatch)
    a = m.arrays()
    tri = a["node_triangle"].copy()
    leaves = np.nonzero(tri >= 0)[0]
    tri[leaves[0]], tri[leaves[1]] = tri[leaves[1]], tri[leaves[0]]
    assert _digest(tri) != case["tri"]
    m.close()
