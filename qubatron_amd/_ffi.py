"""ctypes declarations of the two in-tree libraries.

``libqbt_host.so`` (host C: octree builder with the node numbering of
``src/qubatron/octree.c``, scene generator, flat files) and ``libqbt.so``
(the gfx950 HIP layer whose C ABI replaces ``src/qubatron/octree_glc.c``).
Struct layouts mirror ``include/qbt_host.h`` and ``include/qbt.h``.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")
HOST_LIB = os.path.join(LIBDIR, "libqbt_host.so")
DEV_LIB = os.path.join(LIBDIR, "libqbt.so")

QBT_OK = 0
QBT_EINVAL = -22
QBT_ENOMEM = -12
QBT_ERANGE = -34
QBT_EIO = -5
QBT_EDEVICE = -19
QBT_ETREE = -74
QBT_EABORT = -125
QBT_ENOTSUP = -95

ERRNAMES = {0: "QBT_OK", -22: "QBT_EINVAL", -12: "QBT_ENOMEM", -34: "QBT_ERANGE", -5: "QBT_EIO",
            -19: "QBT_EDEVICE", -74: "QBT_ETREE", -125: "QBT_EABORT", -95: "QBT_ENOTSUP"}

MAX_BOXES, MAX_SPHERES, MAX_ELLIPSOIDS, MAX_LIGHTS = 256, 64, 16, 8

f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)
u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)


class QbtError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        super().__init__(f"{what}: {ERRNAMES.get(code, code)}")


class Octree(C.Structure):
    _fields_ = [("basecube", C.c_float * 4), ("octs", i32p), ("size", C.c_size_t), ("len", C.c_size_t),
                ("levels", C.c_int)]


class SceneSpec(C.Structure):
    _fields_ = [("config", C.c_int), ("seed", C.c_uint64), ("npoints", C.c_uint64), ("depth", C.c_int),
                ("size", C.c_float), ("inset", C.c_float), ("n_planes", C.c_uint64), ("n_objects", C.c_uint64),
                ("n_cluster", C.c_uint64), ("nboxes", C.c_int), ("nspheres", C.c_int), ("nellipsoids", C.c_int),
                ("box_lo", (C.c_float * 3) * MAX_BOXES), ("box_hi", (C.c_float * 3) * MAX_BOXES),
                ("sph_c", (C.c_float * 3) * MAX_SPHERES), ("sph_r", C.c_float * MAX_SPHERES),
                ("ell_c", (C.c_float * 3) * MAX_ELLIPSOIDS), ("ell_a", (C.c_float * 3) * MAX_ELLIPSOIDS),
                ("ell_jitter", C.c_float), ("nlights", C.c_int), ("lights", (C.c_float * 3) * MAX_LIGHTS)]


class Opts(C.Structure):
    _fields_ = [("device", C.c_int), ("max_device_bytes", C.c_size_t)]


class View(C.Structure):
    _fields_ = [("camfp", C.c_float * 3), ("angle", C.c_float * 2), ("width", C.c_float), ("height", C.c_float),
                ("shoot", C.c_int), ("basecube", C.c_float * 4), ("maxlevel", C.c_int)]


class Frame(C.Structure):
    _fields_ = [("w", C.c_int), ("h", C.c_int), ("hit", C.c_void_p), ("leaf", C.c_void_p), ("shadow", C.c_void_p),
                ("rgba8", C.c_void_p), ("rgbaf", C.c_void_p), ("isp", C.c_void_p), ("flags", C.c_void_p),
                ("dyn", C.c_void_p), ("exp", C.c_void_p)]


class Profile(C.Structure):
    _fields_ = [("launches", C.c_uint32), ("reserved", C.c_uint32), ("total_ms", C.c_double), ("avg_ms", C.c_double),
                ("min_ms", C.c_double), ("max_ms", C.c_double), ("kind_launches", C.c_uint32 * 4),
                ("kind_ms", C.c_double * 4)]


class Tuning(C.Structure):
    _fields_ = [("path", C.c_int), ("block_threads", C.c_int), ("refill", C.c_int), ("layout", C.c_int),
                ("stack", C.c_int), ("reorder", C.c_int), ("graph", C.c_int), ("strip", C.c_int),
                ("merge", C.c_int)]


class TuneResult(C.Structure):
    _fields_ = [("tuning", Tuning), ("ms", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("trace_ms", C.c_double), ("primary_rays", C.c_uint64),
                ("shadow_rays", C.c_uint64), ("hits", C.c_uint64), ("expansions", C.c_uint64),
                ("predicted", C.c_uint64), ("launches", C.c_uint32), ("grid_exact", C.c_uint32),
                ("dual", C.c_uint32), ("compact", C.c_uint32), ("expansions_primary", C.c_uint64),
                ("shadow_deferred", C.c_uint64)]


BUFFER_STATIC_COLOR, BUFFER_STATIC_NORMAL, BUFFER_STATIC_OCTREE = 0, 1, 2
BUFFER_DYNAMIC_COLOR, BUFFER_DYNAMIC_NORMAL, BUFFER_DYNAMIC_OCTREE = 3, 4, 5
BUFFER_STATIC_POINTS, BUFFER_DYNAMIC_POINTS = 6, 7
TYPE_INT, TYPE_FLOAT = 0x1404, 0x1406

OUT_HIT, OUT_LEAF, OUT_SHADOW, OUT_RGBA8 = 0x1, 0x2, 0x4, 0x8
OUT_RGBAF, OUT_ISP, OUT_FLAGS, OUT_DYN, OUT_EXP = 0x10, 0x20, 0x40, 0x80, 0x100
OUT_DEFAULT = OUT_HIT | OUT_LEAF | OUT_SHADOW | OUT_RGBA8
OUT_ALL = 0x1FF
RENDER_COUNT = 0x10000
RENDER_SURROGATE = 0x20000
RENDER_COLLECT = 0x40000
MLP_INPUTS = 42
MLP_PARAMS = 11137

PX_DISCARD, PX_LEAF, PX_SHADED, PX_SPRITE, PX_PREDICTED = 1, 2, 4, 8, 16

# exported C-ABI symbols (checked against the headers by the tests)
HOST_SYMBOLS = [
    "qbt_octree_init", "qbt_octree_free", "qbt_octree_reset", "qbt_octree_reserve", "qbt_octet_of",
    "qbt_octree_insert_point", "qbt_octree_insert_path", "qbt_octree_remove_point", "qbt_octree_model_index",
    "qbt_octree_insert_points", "qbt_octree_paths", "qbt_octree_hash", "qbt_flat_count", "qbt_flat_read",
    "qbt_flat_write", "qbt_qmc_quantize", "qbt_scene_preset", "qbt_scene_points", "qbt_scene_blocked",
    "qbt_scene_poses", "qbt_rng_u64", "qbt_rng_f32",
]

_host = None
_dev = None


def _declare_host(h):
    vp = C.c_void_p
    h.qbt_octree_init.argtypes = [C.POINTER(Octree), f32p, C.c_int]
    h.qbt_octree_free.argtypes = [C.POINTER(Octree)]
    h.qbt_octree_free.restype = None
    h.qbt_octree_reset.argtypes = [C.POINTER(Octree)]
    h.qbt_octree_reset.restype = None
    h.qbt_octree_reserve.argtypes = [C.POINTER(Octree), C.c_size_t]
    h.qbt_octet_of.argtypes = [f32p, C.c_float]
    h.qbt_octree_insert_point.argtypes = [C.POINTER(Octree), C.c_size_t, C.c_int32, f32p, i32p, u8p]
    h.qbt_octree_insert_path.argtypes = [C.POINTER(Octree), C.c_size_t, C.c_int32, i32p, i32p, i32p]
    h.qbt_octree_remove_point.argtypes = [C.POINTER(Octree), f32p, i32p, i32p]
    h.qbt_octree_model_index.argtypes = [C.POINTER(Octree), f32p, i32p, i32p]
    h.qbt_octree_insert_points.argtypes = [C.POINTER(Octree), f32p, C.c_size_t, C.c_int32]
    h.qbt_octree_paths.argtypes = [f32p, C.c_int, f32p, C.c_size_t, u8p, C.c_int]
    h.qbt_octree_hash.argtypes = [i32p, C.c_size_t]
    h.qbt_octree_hash.restype = C.c_uint64
    h.qbt_flat_count.argtypes = [C.c_char_p]
    h.qbt_flat_count.restype = C.c_int64
    h.qbt_flat_read.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, f32p, f32p, f32p, C.c_size_t]
    h.qbt_flat_read.restype = C.c_int64
    h.qbt_flat_write.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, f32p, f32p, f32p, C.c_size_t]
    h.qbt_qmc_quantize.argtypes = [f32p, f32p, f32p, C.c_size_t, C.c_float, C.c_int, i32p,
                                   C.POINTER(C.c_size_t)]
    h.qbt_qmc_quantize.restype = C.c_int64
    h.qbt_scene_preset.argtypes = [C.POINTER(SceneSpec), C.c_int, C.c_uint64, C.c_uint64]
    h.qbt_scene_points.argtypes = [C.POINTER(SceneSpec), C.c_uint64, C.c_uint64, f32p, f32p, f32p, C.c_int]
    h.qbt_scene_blocked.argtypes = [C.POINTER(SceneSpec), f32p, C.c_float]
    h.qbt_scene_poses.argtypes = [C.POINTER(SceneSpec), C.c_uint64, C.c_int, C.c_float, f32p, f32p]
    h.qbt_rng_u64.argtypes = [C.c_uint64, C.c_uint64]
    h.qbt_rng_u64.restype = C.c_uint64
    h.qbt_rng_f32.argtypes = [C.c_uint64, C.c_uint64]
    h.qbt_rng_f32.restype = C.c_float
    del vp


def host():
    """The host C library (no GPU needed)."""
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB):
            raise FileNotFoundError(f"{HOST_LIB} is not built: run __graft_entry__.build()")
        h = C.CDLL(HOST_LIB)
        _declare_host(h)
        _host = h
    return _host


DEV_SYMBOLS = [
    "qbt_create", "qbt_destroy", "qbt_strerror", "qbt_device_bytes", "qbt_upload", "qbt_download", "qbt_render",
    "qbt_update", "qbt_read_frame", "qbt_frame_device", "qbt_get_stats", "qbt_trace_rays", "qbt_build_octree",
    "qbt_version", "qbt_last_hip_error", "qbt_synchronize", "qbt_profile_begin", "qbt_profile_end",
    "qbt_copy_bandwidth", "qbt_get_tuning", "qbt_set_tuning",
    "qbt_surrogate_init", "qbt_surrogate_set_params", "qbt_surrogate_get_params", "qbt_surrogate_set_band",
    "qbt_surrogate_reserve", "qbt_surrogate_samples", "qbt_surrogate_read_samples", "qbt_surrogate_forward",
    "qbt_surrogate_predict", "qbt_surrogate_grad", "qbt_surrogate_step", "qbt_surrogate_train", "qbt_autotune",
    "qbt_particle_update", "qbt_render_noalloc", "qbt_layout_stats",
]


def _declare_dev(d):
    vp = C.c_void_p
    d.qbt_create.argtypes = [C.POINTER(Opts), C.POINTER(vp)]
    d.qbt_destroy.argtypes = [vp]
    d.qbt_destroy.restype = None
    d.qbt_strerror.argtypes = [C.c_int]
    d.qbt_strerror.restype = C.c_char_p
    d.qbt_device_bytes.argtypes = [vp]
    d.qbt_device_bytes.restype = C.c_size_t
    d.qbt_upload.argtypes = [vp, vp, C.c_int, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, C.c_int]
    d.qbt_download.argtypes = [vp, C.c_int, vp, C.c_size_t, C.c_size_t]
    d.qbt_render.argtypes = [vp, C.POINTER(View), C.c_int, f32p, C.c_int, C.c_uint]
    d.qbt_render_noalloc.argtypes = [vp, C.POINTER(View), C.c_int, f32p, C.c_int, C.c_uint, C.POINTER(Frame)]
    d.qbt_update.argtypes = [vp, C.c_float, C.c_float, f32p, f32p, C.c_float, C.c_uint8, C.c_int, C.c_float,
                             C.c_int]
    d.qbt_read_frame.argtypes = [vp, C.c_int, C.POINTER(Frame)]
    d.qbt_frame_device.argtypes = [vp, C.c_int, C.POINTER(Frame)]
    d.qbt_get_stats.argtypes = [vp, C.POINTER(Stats)]
    d.qbt_trace_rays.argtypes = [vp, f32p, f32p, C.c_size_t, C.c_int, f32p, C.c_int, i32p, i32p, i32p, f32p,
                                 f32p]
    d.qbt_build_octree.argtypes = [vp, C.c_int, f32p, C.c_int, C.c_size_t, C.POINTER(C.c_size_t)]
    d.qbt_particle_update.argtypes = [vp, f32p, f32p, C.c_size_t, C.c_int, C.c_float, f32p, f32p]
    d.qbt_synchronize.argtypes = [vp]
    u64p = C.POINTER(C.c_uint64)
    d.qbt_layout_stats.argtypes = [vp, C.c_int, u64p, u64p, u64p]
    d.qbt_profile_begin.argtypes = [vp]
    d.qbt_profile_end.argtypes = [vp, C.POINTER(Profile)]
    d.qbt_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
    d.qbt_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    d.qbt_autotune.argtypes = [vp, C.POINTER(View), C.c_int, f32p, C.c_int, C.c_uint, C.c_int,
                               C.POINTER(TuneResult), C.c_int, C.POINTER(C.c_int)]
    d.qbt_copy_bandwidth.argtypes = [vp, C.c_size_t, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    szp = C.POINTER(C.c_size_t)
    d.qbt_surrogate_init.argtypes = [vp, C.c_uint64]
    d.qbt_surrogate_set_params.argtypes = [vp, f32p, C.c_size_t]
    d.qbt_surrogate_get_params.argtypes = [vp, f32p, C.c_size_t]
    d.qbt_surrogate_set_band.argtypes = [vp, C.c_float, C.c_float]
    d.qbt_surrogate_reserve.argtypes = [vp, C.c_size_t, C.c_uint32]
    d.qbt_surrogate_samples.argtypes = [vp, szp, szp]
    d.qbt_surrogate_read_samples.argtypes = [vp, C.c_size_t, C.c_size_t, f32p]
    d.qbt_surrogate_forward.argtypes = [vp, f32p, C.c_size_t, f32p]
    d.qbt_surrogate_predict.argtypes = [vp, f32p, f32p, f32p, C.c_size_t, f32p, f32p]
    d.qbt_surrogate_grad.argtypes = [vp, f32p, f32p, C.c_size_t, f32p, f32p]
    d.qbt_surrogate_step.argtypes = [vp, f32p, f32p, C.c_size_t, C.c_float, f32p]
    d.qbt_surrogate_train.argtypes = [vp, f32p, C.c_int, C.c_int, C.c_float, C.c_uint64, f32p]
    d.qbt_version.restype = C.c_char_p
    d.qbt_last_hip_error.restype = C.c_char_p


def dev():
    """The HIP layer. Raises if the library is missing: there is no CPU fallback."""
    global _dev
    if _dev is None:
        host()  # libqbt.so links libqbt_host.so
        if not os.path.exists(DEV_LIB):
            raise FileNotFoundError(f"{DEV_LIB} is not built: run __graft_entry__.build()")
        d = C.CDLL(DEV_LIB)
        _declare_dev(d)
        _dev = d
    return _dev


def check(rc, what=""):
    if rc < 0:
        extra = ""
        if rc == QBT_EDEVICE and _dev is not None:
            extra = " " + (_dev.qbt_last_hip_error() or b"").decode()
        raise QbtError(rc, what + extra)
    return rc


def fptr(a):
    return a.ctypes.data_as(f32p) if a is not None else None


def iptr(a):
    return a.ctypes.data_as(i32p) if a is not None else None


def u8ptr(a):
    return a.ctypes.data_as(u8p) if a is not None else None
