"""CPU oracle of the SPPM pass loop -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, by __graft_entry__.smoke() and by bench.py's CPU baseline, always as the
checker of the HIP path, never as the thing measured or shipped.  The library is built by
oracle/build.py from oracle.c; the optional reference host build (oracle/_ref) exposes the
reference's own mesh.cpp/bvh.cpp/camera.cpp for cross-checks.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "_build", "libppm_oracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref_host.so")

_lib = None
_ref = None


def library():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError("oracle not built: run python oracle/build.py")
        lib = ctypes.CDLL(ORACLE_LIB)
        vp = ctypes.c_void_p
        fp = ctypes.POINTER(ctypes.c_float)
        lib.oracle_create.restype = ctypes.c_int
        lib.oracle_create.argtypes = [vp, vp, ctypes.POINTER(vp)]
        lib.oracle_pass.restype = ctypes.c_int
        lib.oracle_pass.argtypes = [vp, ctypes.c_uint32]
        lib.oracle_readback.restype = ctypes.c_int
        lib.oracle_readback.argtypes = [vp, ctypes.c_int, vp, ctypes.c_size_t]
        lib.oracle_buffer_bytes.restype = ctypes.c_size_t
        lib.oracle_buffer_bytes.argtypes = [vp, ctypes.c_int]
        lib.oracle_destroy.restype = None
        lib.oracle_destroy.argtypes = [vp]
        for name in ("oracle_sinf", "oracle_cosf", "oracle_expf", "oracle_logf"):
            getattr(lib, name).restype = ctypes.c_float
            getattr(lib, name).argtypes = [ctypes.c_float]
        lib.oracle_powf.restype = ctypes.c_float
        lib.oracle_powf.argtypes = [ctypes.c_float, ctypes.c_float]
        lib.oracle_gpurnd.restype = ctypes.c_float
        lib.oracle_gpurnd.argtypes = [fp]
        lib.oracle_hash.restype = ctypes.c_float
        lib.oracle_hash.argtypes = [fp, ctypes.c_float, ctypes.c_float, ctypes.c_float]
        lib.oracle_seed_grid.restype = None
        lib.oracle_seed_grid.argtypes = [ctypes.POINTER(ctypes.c_uint32), fp, ctypes.c_size_t]
        lib.oracle_xorshift_frand.restype = ctypes.c_float
        lib.oracle_xorshift_frand.argtypes = [ctypes.POINTER(ctypes.c_uint32)]
        lib.oracle_trace.restype = ctypes.c_int
        lib.oracle_trace.argtypes = [vp, fp, fp, ctypes.c_int, ctypes.c_int, fp, fp]
        lib.oracle_radiance.restype = None
        lib.oracle_radiance.argtypes = [fp, fp, ctypes.c_float, fp, fp]
        lib.oracle_num_threads.restype = ctypes.c_int
        lib.oracle_update.restype = None
        lib.oracle_update.argtypes = [fp, fp, fp, fp, ctypes.c_float, fp, ctypes.c_float]
        lib.oracle_texture_sample.restype = None
        lib.oracle_texture_sample.argtypes = [vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, fp]
        lib.oracle_camera_ray.restype = None
        lib.oracle_camera_ray.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, fp, fp,
                                          fp, fp]
        lib.oracle_light_sample.restype = ctypes.c_int
        lib.oracle_light_sample.argtypes = [vp, fp, fp, fp, fp]
        lib.oracle_ibl_sample.restype = None
        lib.oracle_ibl_sample.argtypes = [vp, fp, fp, fp, fp]
        lib.oracle_fresnel.restype = ctypes.c_float
        lib.oracle_fresnel.argtypes = [fp, fp, ctypes.c_float, ctypes.c_float]
        lib.oracle_photon_bounce.restype = ctypes.c_int
        lib.oracle_photon_bounce.argtypes = [vp, ctypes.c_int, fp, fp, fp, fp, fp, fp, fp]
        lib.oracle_glossy_reflect.restype = None
        lib.oracle_glossy_reflect.argtypes = [fp, fp, ctypes.c_float, fp, fp]
        lib.oracle_eye_path.restype = ctypes.c_int
        lib.oracle_eye_path.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, fp, fp]
        _lib = lib
    return _lib


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def sinf(x):
    return np.float32(library().oracle_sinf(float(np.float32(x))))


def cosf(x):
    return np.float32(library().oracle_cosf(float(np.float32(x))))


def expf(x):
    return np.float32(library().oracle_expf(float(np.float32(x))))


def logf(x):
    return np.float32(library().oracle_logf(float(np.float32(x))))


def powf(x, y):
    return np.float32(library().oracle_powf(float(np.float32(x)), float(np.float32(y))))


def gpurnd(state):
    """One GPURnd step; state is a float32[4] array updated in place."""
    assert state.dtype == np.float32 and state.shape == (4,)
    return np.float32(library().oracle_gpurnd(_fptr(state)))


def cell_hash(cell, grid_scale, hash_scale, hash_num):
    c = np.asarray(cell, np.float32)
    return np.float32(library().oracle_hash(_fptr(c), float(np.float32(grid_scale)), float(np.float32(hash_scale)),
                                            float(np.float32(hash_num))))


def seed_grid(state, texels):
    st = (ctypes.c_uint32 * 4)(*state)
    out = np.empty((texels, 4), np.float32)
    library().oracle_seed_grid(st, _fptr(out), texels)
    return out, list(st)


def xorshift_frand(state):
    """Returns (value, new_state)."""
    st = (ctypes.c_uint32 * 4)(*state)
    v = library().oracle_xorshift_frand(st)
    return np.float32(v), list(st)


def trace(scene_ptr, org, direction, cull=False, photon=False):
    """Closest hit of the oracle's BVH walk: (triangle, t, hit position)."""
    o = np.asarray(org, np.float32)
    d = np.asarray(direction, np.float32)
    t = np.zeros(1, np.float32)
    pos = np.zeros(3, np.float32)
    tri = library().oracle_trace(ctypes.cast(scene_ptr, ctypes.c_void_p), _fptr(o), _fptr(d), int(cull), int(photon),
                                 _fptr(t), _fptr(pos))
    return tri, t[0], pos


def texture_sample(scene_ptr, u, v, r):
    """texture3D of the scene's RGB8 volume (GL_LINEAR, GL_REPEAT) as the oracle evaluates it."""
    out = np.zeros(3, np.float32)
    library().oracle_texture_sample(ctypes.cast(scene_ptr, ctypes.c_void_p), float(u), float(v), float(r), _fptr(out))
    return out


def camera_ray(scene_ptr, params, i, j, aa_x, aa_y, rnd):
    """eyeraytrace.fs's camera ray of pixel (i, j): (origin, direction, filter weight, advanced rnd)."""
    r = np.array(rnd, np.float32)
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    f = np.zeros(1, np.float32)
    library().oracle_camera_ray(ctypes.cast(scene_ptr, ctypes.c_void_p), ctypes.cast(ctypes.pointer(params), ctypes.c_void_p),
                                int(i), int(j), float(aa_x), float(aa_y), _fptr(r), _fptr(o), _fptr(d), _fptr(f))
    return o, d, f[0], r


def glossy_reflect(d, n, g, rnd):
    """glossy_reflect (eyeraytrace.fs:370-392, photontrace.fs:258-280): (direction, advanced rnd)."""
    r = np.array(rnd, np.float32)
    out = np.zeros(3, np.float32)
    library().oracle_glossy_reflect(_fptr(np.asarray(d, np.float32)), _fptr(np.asarray(n, np.float32)),
                                    float(np.float32(g)), _fptr(r), _fptr(out))
    return out, r


def eye_path(scene_ptr, params, i, j, aa_x, aa_y, rnd):
    """One eye path of pixel (i, j) (eyeraytrace.fs:396-547): a dict of the last segment's triangle
    (-1: miss), the hit point's position, colour, normal, the last ray direction, the emission
    before averaging, material and t; and the advanced rnd."""
    r = np.array(rnd, np.float32)
    out = np.zeros(17, np.float32)
    tri = library().oracle_eye_path(ctypes.cast(scene_ptr, ctypes.c_void_p),
                                    ctypes.cast(ctypes.pointer(params), ctypes.c_void_p), int(i), int(j),
                                    float(aa_x), float(aa_y), _fptr(r), _fptr(out))
    return dict(tri=int(tri), pos=out[0:3], col=out[3:6], nrm=out[6:9], dir=out[9:12], emission=out[12:15],
                mat=out[15], t=out[16]), r


def light_sample(scene_ptr, rnd):
    """photontrace.fs's emission from the light triangles: (triangle, flux, direction, origin, rnd)."""
    r = np.array(rnd, np.float32)
    f = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    o = np.zeros(3, np.float32)
    tri = library().oracle_light_sample(ctypes.cast(scene_ptr, ctypes.c_void_p), _fptr(r), _fptr(f), _fptr(d), _fptr(o))
    return tri, f, d, o, r


def ibl_sample(scene_ptr, rnd):
    """photontrace.fs's emission from the environment: (flux, direction, origin, rnd)."""
    r = np.array(rnd, np.float32)
    f = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    o = np.zeros(3, np.float32)
    library().oracle_ibl_sample(ctypes.cast(scene_ptr, ctypes.c_void_p), _fptr(r), _fptr(f), _fptr(d), _fptr(o))
    return f, d, o, r


def fresnel(incom, normal, ni, ne):
    return np.float32(library().oracle_fresnel(_fptr(np.asarray(incom, np.float32)), _fptr(np.asarray(normal, np.float32)),
                                               float(ni), float(ne)))


def photon_bounce(scene_ptr, tri, ipos, pnrm, pdir, rnd, flux):
    """photontrace.fs's bounce of a stored photon: (continues, flux, origin, direction, rnd)."""
    r = np.array(rnd, np.float32)
    f = np.array(flux, np.float32)
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    cont = library().oracle_photon_bounce(ctypes.cast(scene_ptr, ctypes.c_void_p), int(tri),
                                          _fptr(np.asarray(ipos, np.float32)), _fptr(np.asarray(pnrm, np.float32)),
                                          _fptr(np.asarray(pdir, np.float32)), _fptr(r), _fptr(f), _fptr(o), _fptr(d))
    return bool(cont), f, o, d, r


def radiance(flux_radius, emission, total_photons):
    fr = np.asarray(flux_radius, np.float32)
    em = np.asarray(emission, np.float32)
    lin = np.zeros(4, np.float32)
    mp = np.zeros(4, np.float32)
    library().oracle_radiance(_fptr(fr), _fptr(em), float(np.float32(total_photons)), _fptr(lin), _fptr(mp))
    return lin, mp


def update(tau, radius, count, phi, m, refl, alpha):
    """The progressive update of one hit point; returns (tau[3], radius, count)."""
    t = np.array(tau, np.float32)
    r = np.array([radius], np.float32)
    n = np.array([count], np.float32)
    library().oracle_update(_fptr(t), _fptr(r), _fptr(n), _fptr(np.array(phi, np.float32)), float(np.float32(m)),
                            _fptr(np.array(refl, np.float32)), float(np.float32(alpha)))
    return t, r[0], n[0]


def num_threads():
    return int(library().oracle_num_threads())


class OracleRenderer:
    """Same buffers and pass semantics as glslppm_amd.Renderer, computed on the CPU."""

    def __init__(self, mesh, params):
        from glslppm_amd import BUFFERS  # buffer table shared with the product's readback API
        self._buffers = BUFFERS
        self._mesh = mesh
        self.params = params
        self._h = ctypes.c_void_p()
        rc = library().oracle_create(ctypes.cast(mesh.scene, ctypes.c_void_p), ctypes.addressof(params),
                                     ctypes.byref(self._h))
        if rc != 0:
            raise RuntimeError("oracle_create failed: %d" % rc)
        self.q = params.image_width * params.image_height
        self.p = params.photon_width * params.photon_height
        self.h = params.hash_width * params.hash_height

    def run(self, npasses=1):
        rc = library().oracle_pass(self._h, int(npasses))
        if rc != 0:
            raise RuntimeError("oracle_pass failed: %d" % rc)
        return self

    def readback(self, name):
        bid, dt, comp, which = self._buffers[name]
        n = {"q": self.q, "p": self.p, "h": self.h}[which]
        out = np.empty((n, comp) if comp > 1 else (n,), dtype=dt)
        rc = library().oracle_readback(self._h, bid, out.ctypes.data, out.nbytes)
        if rc != 0:
            raise RuntimeError("oracle_readback failed: %d" % rc)
        return out

    def counters(self):
        from glslppm_amd import BUF_COUNTERS, Counters
        c = Counters()
        rc = library().oracle_readback(self._h, BUF_COUNTERS, ctypes.addressof(c), ctypes.sizeof(c))
        if rc != 0:
            raise RuntimeError("oracle_readback failed: %d" % rc)
        return {k: getattr(c, k) for k, _ in Counters._fields_ if k != "reserved"}

    def grid(self):
        from glslppm_amd import BUF_GRID, Grid
        g = Grid()
        rc = library().oracle_readback(self._h, BUF_GRID, ctypes.addressof(g), ctypes.sizeof(g))
        if rc != 0:
            raise RuntimeError("oracle_readback failed: %d" % rc)
        return {"bbmin": list(g.bbmin), "bbmax": list(g.bbmax), "initial_radius": g.initial_radius,
                "grid_scale": g.grid_scale, "hash_scale": g.hash_scale}

    def close(self):
        if self._h and self._h.value:
            library().oracle_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ref_library():
    """The reference's own host code (oracle/_ref), or None when it was not built here."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        lib = ctypes.CDLL(REF_LIB)
        fp = ctypes.POINTER(ctypes.c_float)
        ip = ctypes.POINTER(ctypes.c_int)
        lib.ref_load_obj.restype = ctypes.c_int
        lib.ref_load_obj.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_float]
        lib.ref_mesh_from_arrays.restype = ctypes.c_int
        lib.ref_mesh_from_arrays.argtypes = [ctypes.c_int, fp, fp, ip, ctypes.c_int, ip]
        for n in ("ref_num_triangles", "ref_num_materials", "ref_num_lights", "ref_build_bvh"):
            getattr(lib, n).restype = ctypes.c_int
            getattr(lib, n).argtypes = []
        lib.ref_triangles.argtypes = [fp, fp, ip]
        lib.ref_triangles.restype = None
        lib.ref_materials.argtypes = [fp]
        lib.ref_materials.restype = None
        lib.ref_lights.argtypes = [ip, fp]
        lib.ref_lights.restype = ctypes.c_float
        lib.ref_bbox.argtypes = [fp]
        lib.ref_bbox.restype = None
        lib.ref_bvh.argtypes = [fp, ip, ip]
        lib.ref_bvh.restype = None
        lib.ref_camera.argtypes = [fp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_float, fp]
        lib.ref_camera.restype = None
        lib.ref_texcoords.argtypes = [fp]
        lib.ref_texcoords.restype = None
        lib.ref_texture_info.argtypes = [ctypes.c_int, ip, ip]
        lib.ref_texture_info.restype = ctypes.c_int
        lib.ref_texture.argtypes = [ctypes.c_int, ctypes.c_void_p]
        lib.ref_texture.restype = None
        _ref = lib
    return _ref


def ref_mesh_state(lib):
    """Triangles, materials, lights, bbox of the reference's current TMesh."""
    T, M, L = lib.ref_num_triangles(), lib.ref_num_materials(), lib.ref_num_lights()
    pos = np.zeros((T, 9), np.float32)
    nrm = np.zeros((T, 9), np.float32)
    mat = np.zeros(T, np.int32)
    lib.ref_triangles(_fptr(pos), _fptr(nrm), mat.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    mats = np.zeros((M, 6), np.float32)
    lib.ref_materials(_fptr(mats))
    ltri = np.zeros(L, np.int32)
    lcdf = np.zeros(L, np.float32)
    area = lib.ref_lights(ltri.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _fptr(lcdf))
    bbox = np.zeros(6, np.float32)
    lib.ref_bbox(_fptr(bbox))
    return {"positions": pos.reshape(T, 3, 3), "normals": nrm.reshape(T, 3, 3), "material": mat, "materials": mats,
            "light_triangle": ltri, "light_cdf": lcdf, "light_area": np.float32(area), "bbox": bbox}


def ref_bvh(lib):
    N = lib.ref_build_bvh()
    bounds = np.zeros((N, 6), np.float32)
    tri = np.zeros(N, np.int32)
    links = np.zeros((6, N, 2), np.int32)
    lib.ref_bvh(_fptr(bounds), tri.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                links.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return bounds, tri, links
