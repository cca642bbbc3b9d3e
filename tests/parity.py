"""Helpers that compare the HIP renderer with the CPU oracle buffer by buffer."""
import ctypes

import numpy as np

import glslppm_amd as g
import oracle

STATE_BUFFERS = [
    "hit_position", "hit_reflectance", "hit_normal", "hit_triangle", "hit_t", "eye_rng", "flux_radius",
    "emission_count", "photon_position", "photon_flux", "photon_direction", "photon_rng", "photon_state",
    "photon_slot", "slot_count", "slot_rep", "radiance", "tonemapped", "rgba8",
]


def mismatches(a, b):
    """Number of elements whose bytes differ (NaN payloads compared as bits)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.shape != b.shape:
        return -1
    av = a.view(np.uint8).reshape(a.shape[0], -1)
    bv = b.view(np.uint8).reshape(b.shape[0], -1)
    return int(np.any(av != bv, axis=1).sum())


def compare_all(gpu, cpu, names=STATE_BUFFERS):
    out = {}
    for n in names:
        out[n] = mismatches(gpu.readback(n), cpu.readback(n))
    gc, oc = gpu.counters(), cpu.counters()
    out["counters"] = 0 if gc == oc else 1
    gg, og = gpu.grid(), cpu.grid()
    out["grid"] = 0 if np.array_equal(np.float32(sum([gg["bbmin"], gg["bbmax"], [gg["initial_radius"], gg["grid_scale"], gg["hash_scale"]]], [])),
                                      np.float32(sum([og["bbmin"], og["bbmax"], [og["initial_radius"], og["grid_scale"], og["hash_scale"]]], []))) else 1
    return out


def run_pair(mesh, params, passes):
    gpu = g.Renderer(mesh, params)
    cpu = oracle.OracleRenderer(mesh, params)
    gpu.run(passes)
    cpu.run(passes)
    return gpu, cpu


def format_report(rep):
    return ", ".join("%s=%d" % (k, v) for k, v in rep.items() if v)


def check_hash_properties(r, p):
    """Size-independent properties of the hash after a pass: per-slot counts are the histogram of
    the photons' slots (correction), and the representative is the smallest eligible photon index
    of its slot (scatter, with the depth-clip exclusion of the last row's right half)."""
    slot = r.readback("photon_slot")
    count = r.readback("slot_count")
    rep = r.readback("slot_rep")
    H = p.hash_width * p.hash_height
    valid = slot >= 0
    assert np.array_equal(np.bincount(slot[valid], minlength=H).astype(np.uint32), count)
    P = p.photon_width * p.photon_height
    limit = P - p.photon_width // 2
    idx = np.nonzero(valid & (np.arange(P) < limit))[0]
    best = np.full(H, -1, np.int64)
    order = np.argsort(slot[idx], kind="stable")
    s_sorted, i_sorted = slot[idx][order], idx[order]
    first = np.r_[True, s_sorted[1:] != s_sorted[:-1]]
    best[s_sorted[first]] = i_sorted[first]
    assert np.array_equal(best.astype(np.int32), rep)


class ShrunkLeafBoxes:
    """A built mesh's scene with every leaf box shrunk by 1/8 of its extent on each side, as a
    caller may hand in its own tree: the boxes stay nested, and a ray that passes near a
    triangle's edge misses its leaf box, so the reference skips that triangle.  No leaf box is its
    triangle's vertex box, so the walks must read the leaf boxes from the triangle records; the
    vertex boxes would find hits the reference does not."""

    def __init__(self, mesh):
        self._mesh = mesh
        a = mesh.arrays()
        s = g.Scene.from_buffer_copy(mesh.scene.contents)
        b = a["node_bounds"].astype(np.float32)
        leaf = a["node_triangle"] >= 0
        ext = (b[:, 3:] - b[:, :3]) * np.float32(0.125)
        b[leaf, :3] += ext[leaf]
        b[leaf, 3:] -= ext[leaf]
        self._bounds = np.ascontiguousarray(b.reshape(-1))
        s.node_bounds = self._bounds.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        self._scene = s
        self.scene = ctypes.pointer(s)
