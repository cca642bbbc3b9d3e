"""MI355X-native SPPM pass loop of glslppm, driven through the C ABI of include/ppm.h.

The reference runs one stochastic progressive photon mapping pass per GLUT frame in m_display()
(glslppm.cpp:214-536).  Here the same pass order runs as HIP kernels on one GPU:

    mesh = Mesh.generate(SCENE_CORNELL)         # or Mesh.load_obj("cornell_metal.obj", 0.01)
    mesh.set_camera((0, 0, 1.667), (0, 0, 0), 256, 256, 45.0)
    mesh.build()                                # TBVH::Build + flattening (bvh.cpp, gpurt.cpp)
    r = Renderer(mesh, default_params(256))     # ppm_create
    r.run(64)                                   # ppm_pass: 64 x m_display()
    img = r.readback("rgba8")                   # ppm_readback

This module is a thin host mirror: it loads glslppm_amd/_lib/libppm.so and raises PPMError when
the library is missing or the GPU is unavailable -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libppm.so")

SCENE_CORNELL = 0
SCENE_SPHERES = 1
SCENE_TERRAIN = 2

FLAG_NO_DRAW = 1
FLAG_GRAPH = 2
PASSES_UNLIMITED = 0xFFFFFFFF

BRDF = {"light": -1, "diffuse": 0, "mirror": 1, "glass": 2, "plastic": 3, "glossy_mirror": 4,
        "glossy_glass": 5, "glossy_plastic": 6}


class PPMError(RuntimeError):
    pass


class Camera(ctypes.Structure):
    _fields_ = [("origin", ctypes.c_float * 3), ("u", ctypes.c_float * 3), ("v", ctypes.c_float * 3),
                ("w", ctypes.c_float * 3), ("half_width", ctypes.c_float), ("half_height", ctypes.c_float),
                ("distance", ctypes.c_float)]


class Material(ctypes.Structure):
    _fields_ = [("color", ctypes.c_float * 3), ("eta", ctypes.c_float), ("brdf", ctypes.c_int32),
                ("g", ctypes.c_float)]


_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int32)


class Scene(ctypes.Structure):
    _fields_ = [("num_triangles", ctypes.c_int32), ("positions", _fp), ("normals", _fp), ("material", _ip),
                ("num_materials", ctypes.c_int32), ("materials", ctypes.POINTER(Material)),
                ("num_nodes", ctypes.c_int32), ("node_bounds", _fp), ("node_triangle", _ip), ("node_links", _ip),
                ("num_lights", ctypes.c_int32), ("light_triangle", _ip), ("light_cdf", _fp),
                ("light_area", ctypes.c_float), ("bsphere", ctypes.c_float * 4), ("camera", Camera),
                ("texcoords", _fp), ("texture_size", ctypes.c_int32), ("textures", ctypes.POINTER(ctypes.c_uint8))]


class Params(ctypes.Structure):
    _fields_ = [("image_width", ctypes.c_int32), ("image_height", ctypes.c_int32),
                ("photon_width", ctypes.c_int32), ("photon_height", ctypes.c_int32),
                ("hash_width", ctypes.c_int32), ("hash_height", ctypes.c_int32),
                ("max_path_length", ctypes.c_int32), ("alpha", ctypes.c_float),
                ("initial_footprint", ctypes.c_float), ("fixed_radius", ctypes.c_float),
                ("focal_length", ctypes.c_float), ("aperture", ctypes.c_float),
                ("xorshift", ctypes.c_uint32 * 4), ("device", ctypes.c_int32), ("flags", ctypes.c_uint32)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("passes_per_graph", ctypes.c_uint32)]


class Tuning(ctypes.Structure):
    _fields_ = [("trace_block", ctypes.c_int32), ("tile_wave_width", ctypes.c_int32),
                ("passes_per_graph", ctypes.c_uint32), ("walk_stack", ctypes.c_int32),
                ("eye_overlap", ctypes.c_int32), ("ms_per_pass", ctypes.c_float)]

    def update(self, **kw):
        for k, v in kw.items():
            if v is None:
                continue
            if k == "ms_per_pass" or k not in dict(self._fields_):
                raise PPMError("unknown tuning field %s" % k)
            setattr(self, k, int(v))
        return self

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("hbm_budget", ctypes.c_uint64), ("max_jobs", ctypes.c_int32),
                ("passes_per_step", ctypes.c_uint32), ("passes_per_graph", ctypes.c_uint32)]


class JobInfo(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("target_passes", ctypes.c_uint32), ("passes_done", ctypes.c_uint64),
                ("footprint", ctypes.c_uint64), ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_first_pass", ctypes.c_double), ("t_done", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("queued", ctypes.c_int32), ("resident", ctypes.c_int32), ("done", ctypes.c_int32),
                ("failed", ctypes.c_int32), ("device_bytes", ctypes.c_uint64), ("steps", ctypes.c_uint64),
                ("photon_segments", ctypes.c_uint64), ("busy_seconds", ctypes.c_double)]


class Counters(ctypes.Structure):
    _fields_ = [("frame_count", ctypes.c_uint64), ("eye_samples", ctypes.c_uint64),
                ("photons_emitted", ctypes.c_uint64), ("photon_segments", ctypes.c_uint64),
                ("num_photons_f", ctypes.c_float), ("reserved", ctypes.c_uint32)]


class Grid(ctypes.Structure):
    _fields_ = [("bbmin", ctypes.c_float * 3), ("bbmax", ctypes.c_float * 3), ("initial_radius", ctypes.c_float),
                ("grid_scale", ctypes.c_float), ("hash_scale", ctypes.c_float)]


# name -> (ppm_buf id, dtype, components per element, which size: q / p / h)
BUFFERS = {
    "radiance": (0, np.float32, 4, "q"),
    "tonemapped": (1, np.float32, 4, "q"),
    "rgba8": (2, np.uint8, 4, "q"),
    "hit_position": (3, np.float32, 4, "q"),
    "hit_reflectance": (4, np.float32, 4, "q"),
    "hit_normal": (5, np.float32, 4, "q"),
    "hit_triangle": (6, np.int32, 2, "q"),
    "hit_t": (7, np.float32, 1, "q"),
    "eye_rng": (8, np.float32, 4, "q"),
    "flux_radius": (9, np.float32, 4, "q"),
    "emission_count": (10, np.float32, 4, "q"),
    "photon_position": (11, np.float32, 4, "p"),
    "photon_flux": (12, np.float32, 4, "p"),
    "photon_direction": (13, np.float32, 4, "p"),
    "photon_rng": (14, np.float32, 4, "p"),
    "photon_state": (15, np.int32, 3, "p"),
    "photon_slot": (16, np.int32, 1, "p"),
    "slot_count": (17, np.uint32, 1, "h"),
    "slot_rep": (18, np.int32, 1, "h"),
}
KERNELS = ("eye", "clear", "photon", "gather")
PROFILE_KINDS = 0x100  # ppm_profile_enable: time only the kernel kinds whose bits are set
BUF_COUNTERS = 19
BUF_GRID = 20

_lib_handle = None


def library():
    """The loaded libppm.so; raises PPMError if it has not been built."""
    global _lib_handle
    if _lib_handle is not None:
        return _lib_handle
    if not os.path.exists(LIB_PATH):
        raise PPMError("libppm.so is not built (%s); run python -m glslppm_amd.build" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    sig = {
        "ppm_params_default": (None, [ctypes.POINTER(Params), ctypes.c_int32]),
        "ppm_create": (ctypes.c_int, [ctypes.POINTER(Scene), ctypes.POINTER(Params), ctypes.POINTER(vp)]),
        "ppm_pass": (ctypes.c_int, [vp, ctypes.c_uint32]),
        "ppm_pass_static": (ctypes.c_int, [vp, ctypes.c_uint32]),
        "ppm_readback": (ctypes.c_int, [vp, ctypes.c_int, vp, ctypes.c_size_t]),
        "ppm_buffer_bytes": (ctypes.c_size_t, [vp, ctypes.c_int]),
        "ppm_sync": (ctypes.c_int, [vp]),
        "ppm_last_pass_ms": (ctypes.c_float, [vp]),
        "ppm_destroy": (None, [vp]),
        "ppm_profile_enable": (ctypes.c_int, [vp, ctypes.c_int]),
        "ppm_profile_read": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64)]),
        "ppm_last_error": (ctypes.c_char_p, [vp]),
        "ppm_mesh_load_obj": (ctypes.c_int, [ctypes.c_char_p, _fp, ctypes.c_float, ctypes.POINTER(vp)]),
        "ppm_mesh_from_arrays": (ctypes.c_int, [ctypes.c_int32, _fp, _fp, _ip, ctypes.c_int32,
                                                ctypes.POINTER(Material), ctypes.POINTER(vp)]),
        "ppm_mesh_generate": (ctypes.c_int, [ctypes.c_int32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32,
                                             ctypes.POINTER(vp)]),
        "ppm_mesh_set_camera": (ctypes.c_int, [vp, _fp, _fp, ctypes.c_int32, ctypes.c_int32, ctypes.c_float]),
        "ppm_mesh_set_texcoords": (ctypes.c_int, [vp, _fp]),
        "ppm_mesh_set_texture": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, vp]),
        "ppm_mesh_build": (ctypes.c_int, [vp]),
        "ppm_mesh_scene": (ctypes.POINTER(Scene), [vp]),
        "ppm_mesh_num_triangles": (ctypes.c_int32, [vp]),
        "ppm_mesh_destroy": (None, [vp]),
        "ppm_mesh_last_error": (ctypes.c_char_p, []),
        "ppm_get_tuning": (ctypes.c_int, [vp, ctypes.POINTER(Tuning)]),
        "ppm_set_tuning": (ctypes.c_int, [vp, ctypes.POINTER(Tuning)]),
        "ppm_autotune": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.POINTER(Tuning)]),
        "ppm_batch_get_tuning": (ctypes.c_int, [vp, ctypes.POINTER(Tuning)]),
        "ppm_batch_set_tuning": (ctypes.c_int, [vp, ctypes.POINTER(Tuning)]),
        "ppm_batch_autotune": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.POINTER(Tuning)]),
        "ppm_write_png": (ctypes.c_int, [ctypes.c_char_p, vp, ctypes.c_int32, ctypes.c_int32]),
        "ppm_write_pfm": (ctypes.c_int, [ctypes.c_char_p, vp, ctypes.c_int32, ctypes.c_int32]),
        "ppm_batch_config_default": (None, [ctypes.POINTER(BatchConfig)]),
        "ppm_server_config_default": (None, [ctypes.POINTER(ServerConfig)]),
        "ppm_server_create": (ctypes.c_int, [ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]),
        "ppm_server_add_scene": (ctypes.c_int, [vp, ctypes.POINTER(Scene), ctypes.POINTER(ctypes.c_int64)]),
        "ppm_server_remove_scene": (ctypes.c_int, [vp, ctypes.c_int64]),
        "ppm_server_submit": (ctypes.c_int, [vp, ctypes.c_int64, ctypes.POINTER(Camera), ctypes.POINTER(Params),
                                             ctypes.c_uint32, ctypes.POINTER(ctypes.c_int64)]),
        "ppm_server_wait": (ctypes.c_int, [vp, ctypes.c_int64, ctypes.c_double]),
        "ppm_server_info": (ctypes.c_int, [vp, ctypes.c_int64, ctypes.POINTER(JobInfo)]),
        "ppm_server_result": (ctypes.c_int, [vp, ctypes.c_int64, ctypes.c_int, vp, ctypes.c_size_t]),
        "ppm_server_release": (ctypes.c_int, [vp, ctypes.c_int64]),
        "ppm_server_stats_get": (ctypes.c_int, [vp, ctypes.POINTER(ServerStats)]),
        "ppm_server_now": (ctypes.c_double, [vp]),
        "ppm_server_destroy": (None, [vp]),
        "ppm_server_last_error": (ctypes.c_char_p, [vp]),
        "ppm_batch_create": (ctypes.c_int, [ctypes.POINTER(BatchConfig), ctypes.POINTER(vp)]),
        "ppm_batch_add": (ctypes.c_int, [vp, ctypes.POINTER(Scene), ctypes.POINTER(Params), ctypes.c_uint32,
                                         ctypes.POINTER(ctypes.c_int32)]),
        "ppm_batch_remove": (ctypes.c_int, [vp, ctypes.c_int32]),
        "ppm_batch_set_passes": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_uint32]),
        "ppm_batch_pass": (ctypes.c_int, [vp, ctypes.c_uint32]),
        "ppm_batch_sync": (ctypes.c_int, [vp]),
        "ppm_batch_last_pass_ms": (ctypes.c_float, [vp]),
        "ppm_batch_progress": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_uint64),
                                              ctypes.POINTER(ctypes.c_uint32)]),
        "ppm_batch_readback": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int, vp, ctypes.c_size_t]),
        "ppm_batch_buffer_bytes": (ctypes.c_size_t, [vp, ctypes.c_int32, ctypes.c_int]),
        "ppm_batch_num_jobs": (ctypes.c_int32, [vp]),
        "ppm_batch_device_bytes": (ctypes.c_uint64, [vp]),
        "ppm_scene_footprint": (ctypes.c_uint64, [ctypes.POINTER(Scene), ctypes.POINTER(Params)]),
        "ppm_state_footprint": (ctypes.c_uint64, [ctypes.POINTER(Params)]),
        "ppm_batch_destroy": (None, [vp]),
        "ppm_batch_last_error": (ctypes.c_char_p, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib_handle = lib
    return lib


def _f3(v):
    return (ctypes.c_float * 3)(*[float(x) for x in v])


def default_params(resolution=512, **kw):
    """The reference's constants (glslppm.cpp:28-40) at a square size; keywords override fields."""
    p = Params()
    library().ppm_params_default(ctypes.byref(p), int(resolution))
    for k, v in kw.items():
        if k == "xorshift":
            p.xorshift = (ctypes.c_uint32 * 4)(*v)
        elif not hasattr(p, k):
            raise PPMError("unknown parameter %s" % k)
        else:
            setattr(p, k, v)
    return p


class Mesh:
    """Host scene layer: TMesh + TCamera + TBVH of the reference (mesh.cpp, camera.cpp, bvh.cpp)."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle)

    @staticmethod
    def _check(rc):
        if rc != 0:
            raise PPMError("ppm_mesh error %d: %s" % (rc, library().ppm_mesh_last_error().decode()))

    @classmethod
    def generate(cls, kind=SCENE_CORNELL, seed=0x67736C70, count=40, detail=4):
        h = ctypes.c_void_p()
        cls._check(library().ppm_mesh_generate(int(kind), int(seed), int(count), int(detail), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def load_obj(cls, path, scale=1.0, position=(0.0, 0.0, 0.0)):
        h = ctypes.c_void_p()
        cls._check(library().ppm_mesh_load_obj(os.fsencode(path), _f3(position), float(scale), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def from_arrays(cls, positions, materials, material_index=None, normals=None):
        """positions [T,3,3]; materials: list of (rgb, eta, brdf, g); normals [T,3,3] or None."""
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 9)
        T = pos.shape[0]
        mats = (Material * len(materials))()
        for i, (rgb, eta, brdf, g) in enumerate(materials):
            mats[i].color = (ctypes.c_float * 3)(*rgb)
            mats[i].eta, mats[i].brdf, mats[i].g = eta, brdf, g
        mi = np.zeros(T, np.int32) if material_index is None else np.ascontiguousarray(material_index, np.int32)
        nrm = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 9)
        h = ctypes.c_void_p()
        cls._check(library().ppm_mesh_from_arrays(
            T, pos.ctypes.data_as(_fp), None if nrm is None else nrm.ctypes.data_as(_fp), mi.ctypes.data_as(_ip),
            len(materials), mats, ctypes.byref(h)))
        return cls(h.value)

    def set_camera(self, origin, lookat, width, height, fov):
        self._check(library().ppm_mesh_set_camera(self._h, _f3(origin), _f3(lookat), int(width), int(height),
                                                  float(fov)))
        return self

    def build(self):
        self._check(library().ppm_mesh_build(self._h))
        return self

    def set_texcoords(self, uv):
        """Per-vertex texture coordinates [T, 3, 2] (None clears them); rebuild afterwards."""
        if uv is None:
            self._check(library().ppm_mesh_set_texcoords(self._h, None))
        else:
            a = np.ascontiguousarray(uv, dtype=np.float32).reshape(-1)
            self._check(library().ppm_mesh_set_texcoords(self._h, a.ctypes.data_as(_fp)))
        return self

    def set_texture(self, material, image):
        """An RGB8 image [height, width, 3] for a material (None: untextured); rebuild afterwards."""
        if image is None:
            self._check(library().ppm_mesh_set_texture(self._h, int(material), 0, 0, None))
        else:
            a = np.ascontiguousarray(image, dtype=np.uint8)
            self._check(library().ppm_mesh_set_texture(self._h, int(material), a.shape[1], a.shape[0],
                                                       a.ctypes.data))
        return self

    @property
    def num_triangles(self):
        return int(library().ppm_mesh_num_triangles(self._h))

    @property
    def scene(self):
        """Pointer to the flattened ppm_scene (valid after build)."""
        s = library().ppm_mesh_scene(self._h)
        if not s:
            raise PPMError("mesh is not built")
        return s

    def arrays(self):
        """Copies of the flattened scene arrays as numpy arrays."""
        s = self.scene.contents
        T, N, M, L = s.num_triangles, s.num_nodes, s.num_materials, s.num_lights

        def arr(ptr, n, dt):
            if n == 0:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)

        mats = [(tuple(s.materials[i].color), s.materials[i].eta, s.materials[i].brdf, s.materials[i].g)
                for i in range(M)]
        c = s.camera
        return {
            "positions": arr(s.positions, T * 9, np.float32).reshape(T, 3, 3),
            "normals": arr(s.normals, T * 9, np.float32).reshape(T, 3, 3),
            "material": arr(s.material, T, np.int32),
            "materials": mats,
            "node_bounds": arr(s.node_bounds, N * 6, np.float32).reshape(N, 6),
            "node_triangle": arr(s.node_triangle, N, np.int32),
            "node_links": arr(s.node_links, 6 * N * 2, np.int32).reshape(6, N, 2),
            "light_triangle": arr(s.light_triangle, L, np.int32),
            "light_cdf": arr(s.light_cdf, L, np.float32),
            "light_area": np.float32(s.light_area),
            "bsphere": np.array(list(s.bsphere), np.float32),
            "texcoords": (arr(s.texcoords, T * 6, np.float32).reshape(T, 3, 2) if s.texcoords
                          else None),
            "texture_size": int(s.texture_size),
            "textures": (arr(s.textures, M * s.texture_size * s.texture_size * 3, np.uint8).reshape(
                M, s.texture_size, s.texture_size, 3) if s.textures else None),
            "camera": np.array(list(c.origin) + list(c.u) + list(c.v) + list(c.w) +
                               [c.half_width, c.half_height, c.distance], np.float32),
        }

    def close(self):
        if self._h and self._h.value:
            library().ppm_mesh_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Renderer:
    """One SPPM pass loop on one GPU: ppm_create / ppm_pass / ppm_readback."""

    def __init__(self, mesh, params):
        self._mesh = mesh  # keeps the scene arrays alive for the caller's lifetime of this object
        self.params = params
        self._h = ctypes.c_void_p()
        lib = library()
        rc = lib.ppm_create(mesh.scene, ctypes.byref(params), ctypes.byref(self._h))
        if rc != 0:
            raise PPMError("ppm_create failed (%d): %s" % (rc, lib.ppm_last_error(None).decode()))
        self.q = params.image_width * params.image_height
        self.p = params.photon_width * params.photon_height
        self.h = params.hash_width * params.hash_height

    def _check(self, rc):
        if rc != 0:
            raise PPMError("ppm error %d: %s" % (rc, library().ppm_last_error(self._h).decode()))

    def run(self, npasses=1):
        """Queues npasses passes (m_display) on the GPU stream; returns immediately."""
        self._check(library().ppm_pass(self._h, int(npasses)))
        return self

    def run_static(self, npasses=1):
        """ppm_pass_static: the same passes with their partial results in the library's static
        device scratch (no allocation, no graphs); calls on one GPU must not overlap."""
        self._check(library().ppm_pass_static(self._h, int(npasses)))
        return self

    def sync(self):
        self._check(library().ppm_sync(self._h))

    def last_pass_ms(self):
        return float(library().ppm_last_pass_ms(self._h))

    def tuning(self):
        t = Tuning()
        self._check(library().ppm_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, **fields):
        """Changes the named launch-shape fields (trace_block, tile_wave_width, passes_per_graph,
        walk_stack, eye_overlap); the others keep their values."""
        t = Tuning()
        self._check(library().ppm_get_tuning(self._h, ctypes.byref(t)))
        self._check(library().ppm_set_tuning(self._h, ctypes.byref(t.update(**fields))))

    def autotune(self, passes_per_trial=0):
        """Times real passes under each launch shape and keeps the fastest; returns it."""
        t = Tuning()
        self._check(library().ppm_autotune(self._h, int(passes_per_trial), ctypes.byref(t)))
        return t.as_dict()

    def profile(self, enable=True, kinds=None):
        """Starts (and clears) or stops per-kernel event timing; `kinds` (names from KERNELS)
        limits the timing to those kernel kinds."""
        if not enable:
            arg = 0
        elif kinds is None:
            arg = 1
        else:
            arg = PROFILE_KINDS
            for k in kinds:
                arg |= 1 << KERNELS.index(k)
        self._check(library().ppm_profile_enable(self._h, arg))

    def profile_read(self):
        """{kernel: (total ms, launches)} accumulated since profile(True)."""
        ms = (ctypes.c_double * 4)()
        n = (ctypes.c_uint64 * 4)()
        self._check(library().ppm_profile_read(self._h, ms, n))
        return {k: (ms[i], int(n[i])) for i, k in enumerate(KERNELS)}

    def readback(self, name):
        bid, dt, comp, which = BUFFERS[name]
        n = {"q": self.q, "p": self.p, "h": self.h}[which]
        out = np.empty((n, comp) if comp > 1 else (n,), dtype=dt)
        self._check(library().ppm_readback(self._h, bid, out.ctypes.data, out.nbytes))
        return out

    def save(self, path):
        """Writes the current image: .png (tone-mapped RGBA8) or .pfm (linear radiance)."""
        w, h = self.params.image_width, self.params.image_height
        if str(path).lower().endswith(".pfm"):
            write_pfm(path, self.readback("radiance"), w, h)
        else:
            write_png(path, self.readback("rgba8"), w, h)

    def counters(self):
        c = Counters()
        self._check(library().ppm_readback(self._h, BUF_COUNTERS, ctypes.addressof(c), ctypes.sizeof(c)))
        return {k: getattr(c, k) for k, _ in Counters._fields_ if k != "reserved"}

    def grid(self):
        g = Grid()
        self._check(library().ppm_readback(self._h, BUF_GRID, ctypes.addressof(g), ctypes.sizeof(g)))
        return {"bbmin": list(g.bbmin), "bbmax": list(g.bbmax), "initial_radius": g.initial_radius,
                "grid_scale": g.grid_scale, "hash_scale": g.hash_scale}

    def close(self):
        if self._h and self._h.value:
            library().ppm_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_png(path, rgba8, width, height):
    """The RGBA8 readback (rows bottom-up) as a PNG file."""
    a = np.ascontiguousarray(rgba8, dtype=np.uint8)
    if a.size != width * height * 4:
        raise PPMError("rgba8 must hold width x height x 4 bytes")
    rc = library().ppm_write_png(os.fsencode(path), a.ctypes.data, int(width), int(height))
    if rc != 0:
        raise PPMError("ppm_write_png failed (%d)" % rc)


def write_pfm(path, radiance, width, height):
    """The linear radiance readback (float RGBA, rows bottom-up) as a colour PFM file."""
    a = np.ascontiguousarray(radiance, dtype=np.float32)
    if a.size != width * height * 4:
        raise PPMError("radiance must hold width x height x 4 floats")
    rc = library().ppm_write_pfm(os.fsencode(path), a.ctypes.data, int(width), int(height))
    if rc != 0:
        raise PPMError("ppm_write_pfm failed (%d)" % rc)


def scene_footprint(mesh, params):
    """HBM bytes a job of this scene and these parameters allocates (ppm_scene_footprint)."""
    return int(library().ppm_scene_footprint(mesh.scene, ctypes.byref(params)))


def state_footprint(params):
    """HBM bytes of one job's pass state (ppm_state_footprint)."""
    return int(library().ppm_state_footprint(ctypes.byref(params)))


class Batch:
    """Many jobs advanced together, one launch per stage per pass (ppm_batch_*).

    Each job is what a Renderer holds; after n of its passes its buffers equal those of a
    Renderer of the same mesh and parameters after run(n), bit for bit."""

    def __init__(self, device=0, passes_per_graph=0):
        cfg = BatchConfig()
        library().ppm_batch_config_default(ctypes.byref(cfg))
        cfg.device = int(device)
        cfg.passes_per_graph = int(passes_per_graph)
        self._h = ctypes.c_void_p()
        rc = library().ppm_batch_create(ctypes.byref(cfg), ctypes.byref(self._h))
        if rc != 0:
            raise PPMError("ppm_batch_create failed (%d): %s" % (rc, library().ppm_batch_last_error(None).decode()))
        self._meshes = {}
        self._sizes = {}

    def _check(self, rc):
        if rc != 0:
            raise PPMError("ppm_batch error %d: %s" % (rc, library().ppm_batch_last_error(self._h).decode()))

    def add(self, mesh, params, passes=PASSES_UNLIMITED):
        """Adds a job with a budget of `passes` passes; returns its id."""
        j = ctypes.c_int32()
        self._check(library().ppm_batch_add(self._h, mesh.scene, ctypes.byref(params), int(passes), ctypes.byref(j)))
        self._meshes[j.value] = mesh
        self._sizes[j.value] = (params.image_width * params.image_height, params.photon_width * params.photon_height,
                                params.hash_width * params.hash_height)
        return j.value

    def remove(self, job):
        self._check(library().ppm_batch_remove(self._h, int(job)))
        self._meshes.pop(job, None)
        self._sizes.pop(job, None)

    def set_passes(self, job, passes):
        self._check(library().ppm_batch_set_passes(self._h, int(job), int(passes)))

    def run(self, npasses=1):
        self._check(library().ppm_batch_pass(self._h, int(npasses)))
        return self

    def sync(self):
        self._check(library().ppm_batch_sync(self._h))

    def last_pass_ms(self):
        return float(library().ppm_batch_last_pass_ms(self._h))

    def progress(self, job):
        """(passes run, passes left) of a job; waits for the queued passes."""
        f = ctypes.c_uint64()
        left = ctypes.c_uint32()
        self._check(library().ppm_batch_progress(self._h, int(job), ctypes.byref(f), ctypes.byref(left)))
        return int(f.value), int(left.value)

    def tuning(self):
        t = Tuning()
        self._check(library().ppm_batch_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, **fields):
        """Changes the named launch-shape fields (trace_block, tile_wave_width, passes_per_graph,
        walk_stack, eye_overlap); the others keep their values."""
        t = Tuning()
        self._check(library().ppm_batch_get_tuning(self._h, ctypes.byref(t)))
        self._check(library().ppm_batch_set_tuning(self._h, ctypes.byref(t.update(**fields))))

    def autotune(self, passes_per_trial=0):
        """Times real passes under each launch shape and keeps the fastest; returns it."""
        t = Tuning()
        self._check(library().ppm_batch_autotune(self._h, int(passes_per_trial), ctypes.byref(t)))
        return t.as_dict()

    @property
    def num_jobs(self):
        return int(library().ppm_batch_num_jobs(self._h))

    @property
    def device_bytes(self):
        return int(library().ppm_batch_device_bytes(self._h))

    def readback(self, job, name):
        bid, dt, comp, which = BUFFERS[name]
        q, p, h = self._sizes[job]
        n = {"q": q, "p": p, "h": h}[which]
        out = np.empty((n, comp) if comp > 1 else (n,), dtype=dt)
        self._check(library().ppm_batch_readback(self._h, int(job), bid, out.ctypes.data, out.nbytes))
        return out

    def counters(self, job):
        c = Counters()
        self._check(library().ppm_batch_readback(self._h, int(job), BUF_COUNTERS, ctypes.addressof(c),
                                                 ctypes.sizeof(c)))
        return {k: getattr(c, k) for k, _ in Counters._fields_ if k != "reserved"}

    def grid(self, job):
        g = Grid()
        self._check(library().ppm_batch_readback(self._h, int(job), BUF_GRID, ctypes.addressof(g), ctypes.sizeof(g)))
        return {"bbmin": list(g.bbmin), "bbmax": list(g.bbmax), "initial_radius": g.initial_radius,
                "grid_scale": g.grid_scale, "hash_scale": g.hash_scale}

    def close(self):
        if self._h and self._h.value:
            library().ppm_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


JOB_STATES = ("queued", "running", "done", "failed")
SERVER_BUFFERS = ("radiance", "rgba8", "flux_radius", "emission_count")


class Server:
    """Continuous batching of render jobs (ppm_server_*): submit from any thread, wait, read results.

    A job's results equal those of a Renderer of the same scene, camera and parameters after
    run(target_passes), bit for bit."""

    def __init__(self, device=0, hbm_budget=200 << 30, max_jobs=256, passes_per_step=8, passes_per_graph=0):
        cfg = ServerConfig()
        library().ppm_server_config_default(ctypes.byref(cfg))
        cfg.device, cfg.hbm_budget, cfg.max_jobs = int(device), int(hbm_budget), int(max_jobs)
        cfg.passes_per_step, cfg.passes_per_graph = int(passes_per_step), int(passes_per_graph)
        self._h = ctypes.c_void_p()
        rc = library().ppm_server_create(ctypes.byref(cfg), ctypes.byref(self._h))
        if rc != 0:
            raise PPMError("ppm_server_create failed (%d): %s" % (rc, library().ppm_server_last_error(None).decode()))
        self._q = {}

    def _check(self, rc):
        if rc != 0:
            raise PPMError("ppm_server error %d: %s" % (rc, library().ppm_server_last_error(self._h).decode()))

    def add_scene(self, mesh):
        """Registers the mesh's built scene (packed now, uploaded with its first job); returns its id."""
        sid = ctypes.c_int64()
        self._check(library().ppm_server_add_scene(self._h, mesh.scene, ctypes.byref(sid)))
        return sid.value

    def remove_scene(self, scene_id):
        self._check(library().ppm_server_remove_scene(self._h, int(scene_id)))

    def submit(self, scene_id, params, target_passes, camera=None):
        """Queues a job; camera is a Camera or None (the scene's own).  Returns the job id."""
        j = ctypes.c_int64()
        cam = None if camera is None else ctypes.byref(camera)
        self._check(library().ppm_server_submit(self._h, int(scene_id), cam, ctypes.byref(params),
                                                int(target_passes), ctypes.byref(j)))
        self._q[j.value] = params.image_width * params.image_height
        return j.value

    def wait(self, job, timeout=-1.0):
        """True when done, False on time-out; raises PPMError when the job failed."""
        rc = library().ppm_server_wait(self._h, int(job), float(timeout))
        if rc == 1:
            return False
        self._check(rc)
        return True

    def info(self, job):
        i = JobInfo()
        self._check(library().ppm_server_info(self._h, int(job), ctypes.byref(i)))
        d = {k: getattr(i, k) for k, _ in JobInfo._fields_}
        d["state"] = JOB_STATES[d["state"]]
        return d

    def result(self, job, name):
        bid, dt, comp, _ = BUFFERS[name]
        out = np.empty((self._q[job], comp), dtype=dt)
        self._check(library().ppm_server_result(self._h, int(job), bid, out.ctypes.data, out.nbytes))
        return out

    def counters(self, job):
        c = Counters()
        self._check(library().ppm_server_result(self._h, int(job), BUF_COUNTERS, ctypes.addressof(c),
                                                ctypes.sizeof(c)))
        return {k: getattr(c, k) for k, _ in Counters._fields_ if k != "reserved"}

    def release(self, job):
        self._check(library().ppm_server_release(self._h, int(job)))
        self._q.pop(job, None)

    def stats(self):
        st = ServerStats()
        self._check(library().ppm_server_stats_get(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def now(self):
        return float(library().ppm_server_now(self._h))

    def close(self):
        if self._h and self._h.value:
            library().ppm_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_camera(origin, lookat, width, height, fov):
    """TCamera::Set (camera.cpp:4-18) as a Camera structure, through the host scene layer."""
    m = Mesh.generate(SCENE_CORNELL)
    m.set_camera(origin, lookat, width, height, fov).build()
    c = Camera()
    ctypes.pointer(c)[0] = m.scene.contents.camera
    return c
