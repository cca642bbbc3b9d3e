"""Writes tests/golden/reference_bvh_digests.json: SHA-256 digests of the trees that the
reference's own TBVH::Build (bvh.cpp:244-317, compiled from the reference checkout into
oracle/_ref) builds for the seeded sphere scenes at the sizes where the product's threaded build
takes over (above 1024 triangles per range, and above 262,144 triangles for the threaded sorts),
and for the C2 terrain.  The trees themselves are tens of megabytes; their digests let machines
without the reference checkout check the product's trees bit for bit.

Each entry: generator arguments, triangle count, and the digests of node_bounds (float32 [N][6]),
node_triangle (int32 [N]) and node_links (int32 [6][N][2]) as the reference returns them.
Run from the repository root (the C2 tree takes the reference about a minute):
    python tests/golden/make_bvh_digests.py
"""
import ctypes
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import glslppm_amd as g  # noqa: E402
import oracle  # noqa: E402
from oracle import build as ob  # noqa: E402

# (name, kind, seed, count, detail)
CASES = [
    ("spheres-20k", g.SCENE_SPHERES, 0x67736C70, 4, 4),
    ("spheres-c1", g.SCENE_SPHERES, 0x67736C70, 40, 4),
    ("spheres-307k", g.SCENE_SPHERES, 0x67736C71, 60, 4),
    ("terrain-c2", g.SCENE_TERRAIN, 0x67736C72, 8, 1024),
]


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def reference_tree(lib, mesh):
    """The reference's build of a generated mesh's triangles, normals and materials."""
    a = mesh.arrays()
    T = a["positions"].shape[0]
    pos = np.ascontiguousarray(a["positions"].reshape(T, 3, 3), np.float32)
    nrm = np.ascontiguousarray(a["normals"].reshape(T, 3, 3), np.float32)
    mi = np.ascontiguousarray(a["material"], np.int32)
    brdf = np.array([m[2] for m in a["materials"]], np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    lib.ref_mesh_from_arrays(T, oracle._fptr(pos), oracle._fptr(nrm), mi.ctypes.data_as(ip), len(brdf),
                             brdf.ctypes.data_as(ip))
    return oracle.ref_bvh(lib)


def main():
    ob.build_ref()
    lib = oracle.ref_library()
    if lib is None:
        raise SystemExit("the reference checkout is needed to write the digests")
    out = []
    for (name, kind, seed, count, detail) in CASES:
        m = g.Mesh.generate(kind, seed=seed, count=count, detail=detail).build()
        t = time.time()
        b, tri, links = reference_tree(lib, m)
        print("%s: %d triangles, reference build %.1f s" % (name, m.num_triangles, time.time() - t), flush=True)
        out.append(dict(name=name, kind=kind, seed=seed, count=count, detail=detail, triangles=m.num_triangles,
                        nodes=int(b.shape[0]), bounds=digest(b), tri=digest(tri), links=digest(links)))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_bvh_digests.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(path)


if __name__ == "__main__":
    main()
