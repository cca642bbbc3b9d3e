"""Helpers that compare the HIP renderer with the CPU oracle buffer by buffer."""
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
