"""pt_amd: MI355X-native path-tracing back end with the reference renderer's surface.

Python mirror of the two C ABIs (ctypes):
  * ``Scene``  - include/pthost.h: .scn + glTF loading, the packed scene buffer, light alias
                 tables, camera block, .hdr I/O (the reference's src/scene, src/io).
  * ``Tracer`` - include/pthip.h: the device back end replacing the reference's Vulkan
                 PathTracer/PathTraceDispatcher (src/pathtrace/pathtrace.h:43-103,
                 src/pathtrace/ptdispatch.h:24-55).

The device back end has no CPU fallback: constructing a ``Tracer`` without the built
``libpthip.so`` or without a GPU raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")

PT_OK = 0
PT_NO_HIT = 0xFFFFFFFF
_ERRORS = {-1: "PT_EINVAL", -2: "PT_ENOMEM", -3: "PT_EDEVICE", -4: "PT_ESTATE", -5: "PT_ENODEV"}


class PtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (_ERRORS.get(code, str(code)), msg))
        self.code = code


class MeshDesc(C.Structure):
    _fields_ = [("vertex_start", C.c_uint32), ("index_start", C.c_uint32), ("tri_count", C.c_uint32),
                ("vertex_count", C.c_uint32)]


class TextureDesc(C.Structure):
    _fields_ = [("offset", C.c_uint64), ("width", C.c_uint32), ("height", C.c_uint32), ("channels", C.c_uint32),
                ("srgb", C.c_uint32)]


class SceneBlob(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("vertex_count", C.c_uint64), ("indices", C.POINTER(C.c_uint32)),
                ("index_count", C.c_uint64), ("instances", C.c_void_p), ("instance_count", C.c_uint32),
                ("materials", C.c_void_p), ("material_count", C.c_uint32),
                ("light_words", C.POINTER(C.c_uint32)), ("light_word_count", C.c_uint64),
                ("light_count", C.c_uint32), ("mesh_count", C.c_uint32), ("meshes", C.POINTER(MeshDesc)),
                ("instance_mesh", C.POINTER(C.c_uint32)), ("envcol", C.c_float * 4),
                ("textures", C.POINTER(TextureDesc)), ("texture_count", C.c_uint32), ("reserved", C.c_uint32),
                ("texels", C.POINTER(C.c_uint8)), ("texel_bytes", C.c_uint64)]


class Camera(C.Structure):
    _fields_ = [("inv_view", C.c_float * 16), ("inv_proj", C.c_float * 16), ("near_clip", C.c_float),
                ("far_clip", C.c_float), ("focus_dist", C.c_float), ("lens_radius", C.c_float),
                ("exposure", C.c_float)]


class CameraParams(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("lookdir", C.c_float * 3), ("up", C.c_float * 3),
                ("fov_deg", C.c_float), ("near_clip", C.c_float), ("far_clip", C.c_float),
                ("focus_dist", C.c_float), ("lens_radius", C.c_float), ("exposure", C.c_float),
                ("aspect", C.c_float)]


class SceneInfo(C.Structure):
    _fields_ = [("object_variants", C.c_uint32), ("materials", C.c_uint32), ("point_lights", C.c_uint32),
                ("directional_lights", C.c_uint32), ("emissive_instances", C.c_uint32),
                ("instances", C.c_uint32), ("unique_vertices", C.c_uint64), ("unique_triangles", C.c_uint64),
                ("total_triangles", C.c_uint64), ("light_count", C.c_uint32)]


class Ray(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("tmin", C.c_float), ("dir", C.c_float * 3), ("tmax", C.c_float)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("u", C.c_float), ("v", C.c_float), ("instance", C.c_uint32),
                ("primitive", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("paths", C.c_uint64),
                ("render_ms", C.c_double), ("bvh_build_ms", C.c_double), ("bvh_bytes", C.c_uint64),
                ("sample_index", C.c_uint32), ("max_bounces", C.c_uint32), ("width", C.c_uint32),
                ("height", C.c_uint32), ("node_count", C.c_uint32), ("tri_record_count", C.c_uint32),
                ("ext_node_fetches", C.c_uint64), ("ext_tri_fetches", C.c_uint64),
                ("shadow_node_fetches", C.c_uint64), ("shadow_tri_fetches", C.c_uint64),
                ("extend_ms", C.c_double), ("shade_ms", C.c_double), ("shadow_ms", C.c_double),
                ("extend_launches", C.c_uint32), ("shade_launches", C.c_uint32), ("shadow_launches", C.c_uint32),
                ("refill_launches", C.c_uint32), ("refill_ms", C.c_double),
                ("ext_tlas_node_fetches", C.c_uint64), ("ext_instance_entries", C.c_uint64),
                ("shadow_tlas_node_fetches", C.c_uint64), ("shadow_instance_entries", C.c_uint64)]

PT_OPT_PROFILE_KERNELS = 1
PT_OPT_COUNT_FETCHES = 2
PT_OPT_GRAPHS = 3
PT_OPT_BLOCKS_PER_CU = 4
PT_OPT_POOL_SLOTS = 5
PT_OPT_LDS_STACK = 6
PT_OPT_LEAF_SIZE = 7
PT_OPT_SORT_HITS = 8
PT_OPT_NODE_WIDTH = 9
PT_OPT_BVH_BUILDER = 10
PT_OPT_SPILL_DEPTH = 11
PT_OPT_TLAS_OPEN = 12
PT_OPT_TLAS_BOXES = 13
PT_OPT_TLAS_BUILDER = 14


class TuneResult(C.Structure):
    _fields_ = [("blocks_per_cu", C.c_uint32), ("pool_slots", C.c_uint32), ("lds_stack", C.c_uint32),
                ("leaf_size", C.c_uint32), ("node_width", C.c_uint32), ("trials", C.c_uint32),
                ("ns_per_ray_start", C.c_double), ("ns_per_ray_best", C.c_double),
                ("tlas_open", C.c_uint32), ("reserved", C.c_uint32)]
PT_CACHE_FEATURES = 96


class View(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("max_bounces", C.c_uint32), ("reserved", C.c_uint32), ("camera", Camera)]


class CacheConfig(C.Structure):
    _fields_ = [("enable", C.c_uint32), ("query", C.c_uint32), ("train_stride", C.c_uint32),
                ("train_steps", C.c_uint32), ("batch", C.c_uint32), ("learning_rate", C.c_float),
                ("seed", C.c_uint32), ("frame_spp", C.c_uint32), ("continue_prob", C.c_float),
                ("query_vertex", C.c_uint32), ("ema_decay", C.c_float), ("query_max_pdf", C.c_float)]


class CacheStats(C.Structure):
    _fields_ = [("enabled", C.c_uint32), ("query", C.c_uint32), ("steps", C.c_uint32),
                ("last_records", C.c_uint32), ("records_total", C.c_uint64), ("last_loss", C.c_float),
                ("param_count", C.c_uint32), ("bounds_lo", C.c_float * 3), ("bounds_hi", C.c_float * 3)]


HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("instance", "<u4"), ("primitive", "<u4")])
RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("tmin", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")])

_host = None
_hip = None


def host_lib():
    global _host
    if _host is None:
        path = os.path.join(LIB_DIR, "libpthost.so")
        if not os.path.exists(path):
            raise ImportError("libpthost.so is not built: run python -m pt_amd.build")
        lib = C.CDLL(path)
        lib.pth_scene_load.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
        lib.pth_scene_blob.argtypes = [C.c_void_p]
        lib.pth_scene_blob.restype = C.POINTER(SceneBlob)
        lib.pth_scene_info_get.argtypes = [C.c_void_p, C.POINTER(SceneInfo)]
        lib.pth_scene_camera_params.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(CameraParams)]
        lib.pth_scene_free.argtypes = [C.c_void_p]
        lib.pth_camera_block.argtypes = [C.POINTER(CameraParams), C.POINTER(Camera)]
        lib.pth_write_hdr.argtypes = [C.c_char_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_size_t]
        lib.pth_write_png.argtypes = [C.c_char_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_size_t]
        lib.pth_read_hdr.argtypes = [C.c_char_p, C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_uint32),
                                     C.POINTER(C.c_uint32)]
        lib.pth_meansq.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32]
        lib.pth_meansq.restype = C.c_double
        lib.pth_from_srgb.argtypes = [C.c_float]
        lib.pth_from_srgb.restype = C.c_float
        lib.pth_to_srgb.argtypes = [C.c_float]
        lib.pth_to_srgb.restype = C.c_float
        lib.pth_free.argtypes = [C.c_void_p]
        lib.pth_decode_image.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                         C.POINTER(C.c_uint32), C.POINTER(C.POINTER(C.c_uint8))]
        lib.pth_alias_table.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p]
        lib.pth_last_error.restype = C.c_char_p
        _host = lib
    return _host


def hip_lib():
    """Loads libpthip.so (or the build named by $PT_LIBPTHIP, for comparing builds); raises if it
    is not built (there is no CPU fallback)."""
    global _hip
    if _hip is None:
        path = os.environ.get("PT_LIBPTHIP") or os.path.join(LIB_DIR, "libpthip.so")
        if not os.path.exists(path):
            raise ImportError("libpthip.so is not built: run python -m pt_amd.build")
        lib = C.CDLL(path)
        P = C.c_void_p
        lib.pt_create.argtypes = [C.POINTER(SceneBlob), C.c_uint32, C.c_uint32, C.POINTER(P)]
        lib.pt_set_camera.argtypes = [P, C.POINTER(Camera)]
        lib.pt_set_max_bounces.argtypes = [P, C.c_uint32]
        lib.pt_render.argtypes = [P, C.c_uint32]
        lib.pt_read_accumulation.argtypes = [P, C.c_void_p, C.c_size_t]
        lib.pt_read_tonemapped.argtypes = [P, C.c_void_p, C.c_size_t]
        lib.pt_trace_primary.argtypes = [P, C.c_uint32, C.c_void_p]
        lib.pt_trace_rays.argtypes = [P, C.c_void_p, C.c_uint32, C.c_void_p, C.c_int]
        lib.pt_trace_rays_noalloc.argtypes = [P, C.c_void_p, C.c_uint32, C.c_void_p, C.c_int]
        lib.pt_get_stats.argtypes = [P, C.POINTER(Stats)]
        lib.pt_synchronize.argtypes = [P]
        lib.pt_set_option.argtypes = [P, C.c_uint32, C.c_int64]
        lib.pt_reset_stats.argtypes = [P]
        lib.pt_last_error.argtypes = [P]
        lib.pt_last_error.restype = C.c_char_p
        lib.pt_destroy.argtypes = [P]
        lib.pt_autotune.argtypes = [P, C.c_uint32, C.POINTER(TuneResult)]
        lib.pt_build_light_tables.argtypes = [P]
        lib.pt_read_light_words.argtypes = [P, C.c_void_p, C.c_size_t]
        lib.pt_batch_create.argtypes = [C.POINTER(P), C.c_uint32, C.POINTER(View), C.c_uint32, C.POINTER(P)]
        lib.pt_batch_set_view.argtypes = [P, C.c_uint32, C.POINTER(View)]
        lib.pt_batch_render.argtypes = [P]
        lib.pt_batch_read.argtypes = [P, C.c_uint32, C.c_void_p, C.c_size_t]
        lib.pt_batch_read_tonemapped.argtypes = [P, C.c_uint32, C.c_void_p, C.c_size_t]
        lib.pt_batch_get_stats.argtypes = [P, C.POINTER(Stats)]
        lib.pt_batch_set_option.argtypes = [P, C.c_uint32, C.c_int64]
        lib.pt_batch_reset_stats.argtypes = [P]
        lib.pt_batch_last_error.argtypes = [P]
        lib.pt_batch_last_error.restype = C.c_char_p
        lib.pt_batch_destroy.argtypes = [P]
        lib.pt_cache_configure.argtypes = [P, C.POINTER(CacheConfig)]
        lib.pt_cache_get_stats.argtypes = [P, C.POINTER(CacheStats)]
        lib.pt_cache_param_count.argtypes = []
        lib.pt_cache_param_count.restype = C.c_uint32
        lib.pt_cache_get_params.argtypes = [P, C.c_void_p, C.c_size_t]
        lib.pt_cache_get_query_params.argtypes = [P, C.c_void_p, C.c_size_t]
        lib.pt_cache_set_params.argtypes = [P, C.c_void_p, C.c_size_t]
        lib.pt_cache_forward.argtypes = [P, C.c_void_p, C.c_uint32, C.c_void_p]
        lib.pt_cache_encode.argtypes = [P, C.c_void_p, C.c_uint32, C.c_void_p]
        lib.pt_cache_train_batch.argtypes = [P, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p,
                                             C.c_void_p]
        _hip = lib
    return _hip


def _check_host(code):
    if code != PT_OK:
        raise PtError(code, host_lib().pth_last_error().decode())


def alias_table(weights, values, value_words=1):
    """AliasTable::build + copy: returns the packed words (uint32)."""
    w = np.ascontiguousarray(weights, dtype=np.float32)
    v = np.ascontiguousarray(values, dtype=np.uint32)
    out = np.zeros(len(w) * (3 + value_words), np.uint32)
    _check_host(host_lib().pth_alias_table(w.ctypes.data, len(w), v.ctypes.data, value_words, out.ctypes.data))
    return out


def from_srgb(x):
    return host_lib().pth_from_srgb(x)


def to_srgb(x):
    return host_lib().pth_to_srgb(x)


def camera_block(position, lookdir, up=(0, 1, 0), fov_deg=45.0, aspect=1.0, near=0.1, far=1000.0,
                 focus_dist=0.0, lens_radius=0.0, exposure=1.0):
    p = CameraParams()
    p.position[:] = list(position)
    p.lookdir[:] = list(lookdir)
    p.up[:] = list(up)
    p.fov_deg, p.aspect, p.near_clip, p.far_clip = fov_deg, aspect, near, far
    p.focus_dist, p.lens_radius, p.exposure = focus_dist, lens_radius, exposure
    cam = Camera()
    _check_host(host_lib().pth_camera_block(C.byref(p), C.byref(cam)))
    return cam


class Scene:
    """A loaded .scn scene and its packed buffers (SceneBuilder + SceneBuffer)."""

    def __init__(self, scn_path):
        lib = host_lib()
        h = C.c_void_p()
        _check_host(lib.pth_scene_load(os.fsencode(scn_path), C.byref(h)))
        self._h = h
        self.path = scn_path

    def close(self):
        if getattr(self, "_h", None):
            host_lib().pth_scene_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def blob(self):
        return host_lib().pth_scene_blob(self._h)

    def info(self):
        out = SceneInfo()
        _check_host(host_lib().pth_scene_info_get(self._h, C.byref(out)))
        return {k: getattr(out, k) for k, _ in SceneInfo._fields_}

    def camera_params(self, width, height):
        p = CameraParams()
        _check_host(host_lib().pth_scene_camera_params(self._h, width, height, C.byref(p)))
        return p

    def camera(self, width, height):
        p = self.camera_params(width, height)
        cam = Camera()
        _check_host(host_lib().pth_camera_block(C.byref(p), C.byref(cam)))
        return cam

    # views of the packed arrays (copies)
    def arrays(self):
        b = self.blob.contents
        nv, ni, nin = b.vertex_count, b.index_count, b.instance_count
        out = {
            "vertices": np.ctypeslib.as_array(C.cast(b.vertices, C.POINTER(C.c_float)), (nv * 8,)).reshape(nv, 8).copy(),
            "indices": np.ctypeslib.as_array(b.indices, (ni,)).copy() if ni else np.zeros(0, np.uint32),
            "instances": np.ctypeslib.as_array(C.cast(b.instances, C.POINTER(C.c_uint8)), (nin * 128,)).copy(),
            "materials": np.ctypeslib.as_array(C.cast(b.materials, C.POINTER(C.c_float)),
                                               (b.material_count * 24,)).reshape(-1, 24).copy(),
            "light_words": (np.ctypeslib.as_array(b.light_words, (b.light_word_count,)).copy()
                            if b.light_word_count else np.zeros(0, np.uint32)),
            "light_count": b.light_count,
            "meshes": [(m.vertex_start, m.index_start, m.tri_count, m.vertex_count) for m in b.meshes[:b.mesh_count]],
            "instance_mesh": np.ctypeslib.as_array(b.instance_mesh, (nin,)).copy() if nin else np.zeros(0, np.uint32),
            "envcol": list(b.envcol),
        }
        return out


def write_hdr(path, rgba):
    rgba = np.ascontiguousarray(rgba, dtype=np.float32)
    h, w = rgba.shape[:2]
    _check_host(host_lib().pth_write_hdr(os.fsencode(path), rgba.ctypes.data, w, h, w * 16))


def write_png(path, rgba8):
    """write_png (src/io/imagewrite.cpp:44-67) of H x W x 4 uint8 pixels."""
    rgba8 = np.ascontiguousarray(rgba8, dtype=np.uint8)
    h, w = rgba8.shape[:2]
    _check_host(host_lib().pth_write_png(os.fsencode(path), rgba8.ctypes.data, w, h, w * 4))


def read_hdr(path):
    lib = host_lib()
    ptr = C.POINTER(C.c_float)()
    w, h = C.c_uint32(), C.c_uint32()
    _check_host(lib.pth_read_hdr(os.fsencode(path), C.byref(ptr), C.byref(w), C.byref(h)))
    try:
        return np.ctypeslib.as_array(ptr, (h.value, w.value, 4)).copy()
    finally:
        lib.pth_free(ptr)


def meansq(img, ref):
    img = np.ascontiguousarray(img, dtype=np.float32)
    ref = np.ascontiguousarray(ref, dtype=np.float32)
    assert img.shape == ref.shape
    h, w = img.shape[:2]
    return host_lib().pth_meansq(img.ctypes.data, ref.ctypes.data, w, h)


def rel_mse(img, ref, eps=1e-2):
    """mean((x - r)^2 / (r^2 + eps)) over RGB."""
    d = img[..., :3].astype(np.float64) - ref[..., :3]
    return float(np.mean(d * d / (ref[..., :3].astype(np.float64) ** 2 + eps)))


class Tracer:
    """The device back end (include/pthip.h) for one scene and image size."""

    def __init__(self, scene, width, height, max_bounces=2):
        lib = hip_lib()
        self._lib = lib
        self.scene = scene  # keep the host blob alive until pt_create has copied it
        self.width, self.height = width, height
        h = C.c_void_p()
        code = lib.pt_create(scene.blob, width, height, C.byref(h))
        if code != PT_OK:
            raise PtError(code, lib.pt_last_error(None).decode())
        self._h = h
        self.set_max_bounces(max_bounces)

    def _check(self, code):
        if code != PT_OK:
            raise PtError(code, self._lib.pt_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_camera(self, cam):
        self._check(self._lib.pt_set_camera(self._h, C.byref(cam)))

    def set_max_bounces(self, b):
        self._check(self._lib.pt_set_max_bounces(self._h, b))

    def render(self, spp):
        self._check(self._lib.pt_render(self._h, spp))

    def reset_stats(self):
        self._check(self._lib.pt_reset_stats(self._h))

    def set_option(self, option, value):
        self._check(self._lib.pt_set_option(self._h, option, int(value)))

    def synchronize(self):
        self._check(self._lib.pt_synchronize(self._h))

    def read(self):
        out = np.zeros((self.height, self.width, 4), np.float32)
        self._check(self._lib.pt_read_accumulation(self._h, out.ctypes.data, self.width * 16))
        return out

    def read_tonemapped(self):
        """The PNG output path: ACES tone mapping and 8-bit conversion, H x W x 4 uint8."""
        out = np.zeros((self.height, self.width, 4), np.uint8)
        self._check(self._lib.pt_read_tonemapped(self._h, out.ctypes.data, self.width * 4))
        return out

    def trace_primary(self, sample_index):
        out = np.zeros(self.width * self.height, HIT_DTYPE)
        self._check(self._lib.pt_trace_primary(self._h, sample_index, out.ctypes.data))
        return out.reshape(self.height, self.width)

    def trace_rays(self, rays, any_hit=False):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        out = np.zeros(len(rays), HIT_DTYPE)
        self._check(self._lib.pt_trace_rays(self._h, rays.ctypes.data, len(rays), out.ctypes.data, int(any_hit)))
        return out

    def trace_rays_noalloc(self, rays, any_hit=False):
        """trace_rays through the library's fixed staging array (no device allocation)."""
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        out = np.zeros(len(rays), HIT_DTYPE)
        self._check(self._lib.pt_trace_rays_noalloc(self._h, rays.ctypes.data, len(rays), out.ctypes.data,
                                                    int(any_hit)))
        return out

    def stats(self):
        s = Stats()
        self._check(self._lib.pt_get_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def build_light_tables(self):
        self._check(self._lib.pt_build_light_tables(self._h))

    def light_words(self):
        n = int(self.scene.blob.contents.light_word_count)
        out = np.zeros(n, np.uint32)
        self._check(self._lib.pt_read_light_words(self._h, out.ctypes.data, n))
        return out

    def autotune(self, spp=2):
        r = TuneResult()
        self._check(self._lib.pt_autotune(self._h, spp, C.byref(r)))
        return {k: getattr(r, k) for k, _ in TuneResult._fields_}

    # ---- learned radiance cache (no counterpart in the reference)

    def cache_configure(self, enable=True, query=True, train_stride=32, train_steps=4, batch=16384,
                        learning_rate=1e-3, seed=1, frame_spp=1, continue_prob=0.0, query_vertex=0, ema_decay=0.0,
                        query_max_pdf=0.0):
        cfg = CacheConfig(int(enable), int(query), train_stride, train_steps, batch, learning_rate, seed, frame_spp,
                          continue_prob, query_vertex, ema_decay, query_max_pdf)
        self._check(self._lib.pt_cache_configure(self._h, C.byref(cfg)))

    def cache_stats(self):
        s = CacheStats()
        self._check(self._lib.pt_cache_get_stats(self._h, C.byref(s)))
        out = {k: getattr(s, k) for k, _ in CacheStats._fields_}
        out["bounds_lo"] = list(s.bounds_lo)
        out["bounds_hi"] = list(s.bounds_hi)
        return out

    def cache_params(self):
        n = self._lib.pt_cache_param_count()
        out = np.zeros(n, np.float32)
        self._check(self._lib.pt_cache_get_params(self._h, out.ctypes.data, n))
        return out

    def cache_query_params(self):
        n = self._lib.pt_cache_param_count()
        out = np.zeros(n, np.float32)
        self._check(self._lib.pt_cache_get_query_params(self._h, out.ctypes.data, n))
        return out

    def cache_set_params(self, params):
        params = np.ascontiguousarray(params, np.float32)
        self._check(self._lib.pt_cache_set_params(self._h, params.ctypes.data, params.size))

    def cache_forward(self, features):
        f = np.ascontiguousarray(features, np.float32).reshape(-1, PT_CACHE_FEATURES)
        out = np.zeros((len(f), 3), np.float32)
        self._check(self._lib.pt_cache_forward(self._h, f.ctypes.data, len(f), out.ctypes.data))
        return out

    def cache_encode(self, records):
        r = np.ascontiguousarray(records, np.float32).reshape(-1, 13)
        out = np.zeros((len(r), PT_CACHE_FEATURES), np.float32)
        self._check(self._lib.pt_cache_encode(self._h, r.ctypes.data, len(r), out.ctypes.data))
        return out

    def cache_train_batch(self, features, targets, apply=False):
        """Loss and parameter gradient of one batch; apply=True also takes an Adam step."""
        f = np.ascontiguousarray(features, np.float32).reshape(-1, PT_CACHE_FEATURES)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        grads = np.zeros(self._lib.pt_cache_param_count(), np.float32)
        loss = np.zeros(1, np.float32)
        self._check(self._lib.pt_cache_train_batch(self._h, f.ctypes.data, t.ctypes.data, len(f), int(apply),
                                                   grads.ctypes.data, loss.ctypes.data))
        return float(loss[0]), grads


class Batch:
    """Many views of one or more scenes rendered in one persistent wavefront run
    (pt_batch_* in include/pthip.h). tracers: the Tracer objects whose scenes the views use
    (kept alive by the batch); views: (scene index, width, height, spp, max_bounces, Camera)."""

    def __init__(self, tracers, views):
        lib = hip_lib()
        self._lib = lib
        self.tracers = list(tracers)
        self.views = [tuple(v) for v in views]
        arr = (C.c_void_p * len(self.tracers))(*[t._h.value for t in self.tracers])
        varr = (View * len(self.views))()
        for i, v in enumerate(self.views):
            varr[i] = self._view(v)
        h = C.c_void_p()
        code = lib.pt_batch_create(C.cast(arr, C.POINTER(C.c_void_p)), len(self.tracers), varr, len(self.views),
                                   C.byref(h))
        if code != PT_OK:
            raise PtError(code, lib.pt_batch_last_error(None).decode())
        self._h = h

    @staticmethod
    def _view(v):
        scene, w, h, spp, bounces, cam = v
        return View(scene, w, h, spp, bounces, 0, cam)

    def _check(self, code):
        if code != PT_OK:
            raise PtError(code, self._lib.pt_batch_last_error(self._h).decode())

    def render(self):
        self._check(self._lib.pt_batch_render(self._h))

    def set_view(self, i, v):
        vv = self._view(v)
        self._check(self._lib.pt_batch_set_view(self._h, i, C.byref(vv)))
        self.views[i] = tuple(v)

    def read(self, i):
        _, w, h, _, _, _ = self.views[i]
        out = np.zeros((h, w, 4), np.float32)
        self._check(self._lib.pt_batch_read(self._h, i, out.ctypes.data, w * 16))
        return out

    def read_tonemapped(self, i):
        _, w, h, _, _, _ = self.views[i]
        out = np.zeros((h, w, 4), np.uint8)
        self._check(self._lib.pt_batch_read_tonemapped(self._h, i, out.ctypes.data, w * 4))
        return out

    def stats(self):
        s = Stats()
        self._check(self._lib.pt_batch_get_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def set_option(self, option, value):
        self._check(self._lib.pt_batch_set_option(self._h, option, int(value)))

    def reset_stats(self):
        self._check(self._lib.pt_batch_reset_stats(self._h))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pt_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
