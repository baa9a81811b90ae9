"""raydar_amd — MI355X (gfx950) HIP path-tracing backend for raydar.

The product is the C-ABI library libraydar_hip.so (include/raydar_hip.h);
this package loads it with ctypes and mirrors raydar's `trait Renderer`
(src/renderer/mod.rs:25-35) as `HipRenderer` for Python callers and tests.
There is no CPU fallback: without the built library or without a HIP device
every entry point raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .scene import Scene  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("RAYDAR_HIP_LIB") or os.path.join(_HERE, "libraydar_hip.so")
_lib = None


class RaydarError(RuntimeError):
    def __init__(self, status, message):
        self.status = status
        super().__init__(f"{abi.STATUS_NAMES.get(status, status)}: {message}")


def lib():
    """Load libraydar_hip.so (built in-tree by raydar_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m raydar_amd.build` first")
    L = C.CDLL(LIB_PATH)
    R = C.c_void_p
    P = C.POINTER
    sp = P(abi.RdScene)
    sig = {
        "rd_config_default": (C.c_int, [P(abi.RdConfig)]),
        "rd_create": (C.c_int, [P(abi.RdConfig), P(R)]),
        "rd_destroy": (None, [R]),
        "rd_new_frame": (C.c_int, [R, sp]),
        "rd_render_sample": (C.c_int, [R, sp, C.c_void_p, C.c_size_t, P(C.c_int)]),
        "rd_render_frame": (C.c_int, [R, sp, C.c_void_p, C.c_size_t]),
        "rd_profiler": (C.c_int, [R, P(abi.RdTimings)]),
        "rd_sample_count": (C.c_uint32, [R]),
        "rd_max_sample_count": (C.c_uint32, [R]),
        "rd_max_bounces": (C.c_uint32, [R]),
        "rd_set_max_sample_count": (None, [R, C.c_uint32]),
        "rd_set_max_bounces": (None, [R, C.c_uint32]),
        "rd_last_error": (C.c_char_p, [R]),
        "rd_set_samples_per_pass": (C.c_int, [R, C.c_uint32]),
        "rd_set_seed": (None, [R, C.c_uint64]),
        "rd_read_accumulator": (C.c_int, [R, C.c_void_p, C.c_size_t]),
        "rd_restart_frame": (C.c_int, [R]),
        "rd_set_option": (C.c_int, [R, C.c_char_p, C.c_int64]),
        "rd_get_option": (C.c_int, [R, C.c_char_p, P(C.c_int64)]),
        "rd_get_counters": (C.c_int, [R, P(abi.RdCounters)]),
        "rd_reset_counters": (C.c_int, [R]),
        "rd_get_kernel_times": (C.c_int, [R, P(abi.RdKernelTimes)]),
        "rd_camera_update_matrices": (C.c_int, [P(abi.RdCamera)]),
        "rd_camera_vulkan_inverse_view": (C.c_int, [P(abi.RdCamera), C.c_void_p]),
        "rd_scene_load_rscn": (C.c_int, [C.c_char_p, P(sp)]),
        "rd_scene_parse_rscn": (C.c_int, [C.c_char_p, C.c_size_t, P(sp)]),
        "rd_scene_save_rscn": (C.c_int, [sp, C.c_char_p]),
        "rd_scene_free": (None, [sp]),
        "rd_debug_trace": (C.c_int, [sp, C.c_int, C.c_int, C.c_void_p, C.c_uint32, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p]),
        "rd_debug_extend": (C.c_int, [sp, C.c_int, P(abi.RdExtendVariant), C.c_void_p, C.c_uint32,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
        "rd_debug_scatter": (C.c_int, [sp, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
        "rd_debug_camera_rays": (C.c_int, [sp, C.c_void_p, C.c_uint32, C.c_void_p]),
        "rd_debug_world": (C.c_int, [sp, C.c_void_p, C.c_uint32, C.c_void_p]),
        "rd_debug_philox": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
        "rd_debug_copy_calibration": (C.c_int, [C.c_uint64, C.c_uint32, C.c_void_p]),
        "rd_debug_host_bvh_info": (C.c_int, [sp, C.c_int, P(C.c_uint32), P(C.c_uint32),
                                             C.c_char_p, C.c_size_t]),
        "rd_debug_host_trace": (C.c_int, [sp, C.c_int, C.c_void_p, C.c_uint32, C.c_void_p,
                                          C.c_void_p]),
        "rd_debug_ring_ops": (C.c_int, [C.c_uint64, C.c_void_p, C.c_uint32, C.c_void_p]),
        "rd_debug_cache_classes": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
        "rd_debug_host_bvh4_info": (C.c_int, [sp, C.c_int, P(C.c_uint32), P(C.c_uint32),
                                              C.c_char_p, C.c_size_t]),
        "rd_debug_host_trace4": (C.c_int, [sp, C.c_int, C.c_void_p, C.c_uint32, C.c_void_p,
                                           C.c_void_p]),
        "rd_debug_host_scatter": (C.c_int, [sp, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
        "rd_debug_host_camera_rays": (C.c_int, [sp, C.c_void_p, C.c_uint32, C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def last_error(handle=None):
    msg = lib().rd_last_error(handle)
    return msg.decode() if msg else ""


def check(status, handle=None):
    if status != abi.RD_OK:
        raise RaydarError(status, last_error(handle))


# Random words per bounce in the scatter hooks (two Philox-4x32 blocks).
DRAW_WORDS = 8


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class HipRenderer:
    """Python mirror of raydar's Renderer trait over the C ABI."""

    def __init__(self, max_sample_count=1024, max_bounces=12, seed=0, **options):
        L = lib()
        cfg = abi.RdConfig(max_sample_count, max_bounces)
        h = C.c_void_p()
        check(L.rd_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._scene = None
        L.rd_set_seed(h, seed)
        for k, v in options.items():
            self.set_option(k, v)

    def close(self):
        if self._h:
            lib().rd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _rd(self, scene):
        if scene is None:
            return None
        self._scene = scene  # keep the object array alive
        self._rd_scene = scene.as_rd()
        return C.byref(self._rd_scene)

    # -- trait Renderer ---------------------------------------------------
    def new_frame(self, scene):
        check(lib().rd_new_frame(self._h, self._rd(scene)), self._h)

    @staticmethod
    def _image(sc, out):
        """`out` if given (reused, e.g. a caller's frame buffer), else a new image."""
        if out is None:
            return np.zeros((sc.height, sc.width, 4), np.uint8)
        if out.dtype != np.uint8 or out.size != sc.height * sc.width * 4 or not out.flags.c_contiguous:
            raise ValueError("out must be a C-contiguous uint8 array of height*width*4 elements")
        return out

    def render_frame(self, scene, out=None):
        out = self._image(scene, out)
        check(lib().rd_render_frame(self._h, self._rd(scene), _ptr(out), out.nbytes), self._h)
        return out

    def render_sample(self, scene=None, out=None):
        sc = scene if scene is not None else self._scene
        out = self._image(sc, out)
        produced = C.c_int(0)
        check(lib().rd_render_sample(self._h, self._rd(scene), _ptr(out), out.nbytes,
                                     C.byref(produced)), self._h)
        return out if produced.value else None

    def profiler(self):
        t = abi.RdTimings()
        check(lib().rd_profiler(self._h, C.byref(t)), self._h)
        return {k: (None if getattr(t, k) < 0 else getattr(t, k) / 1e9)
                for k in ("frame_ns", "sample_ns", "prepare_ns", "render_ns")}

    @property
    def sample_count(self):
        return lib().rd_sample_count(self._h)

    @property
    def max_sample_count(self):
        return lib().rd_max_sample_count(self._h)

    @max_sample_count.setter
    def max_sample_count(self, v):
        lib().rd_set_max_sample_count(self._h, v)

    @property
    def max_bounces(self):
        return lib().rd_max_bounces(self._h)

    @max_bounces.setter
    def max_bounces(self, v):
        lib().rd_set_max_bounces(self._h, v)

    # -- extensions -------------------------------------------------------
    def restart_frame(self):
        check(lib().rd_restart_frame(self._h), self._h)

    def set_seed(self, seed):
        lib().rd_set_seed(self._h, seed)

    def set_samples_per_pass(self, n):
        check(lib().rd_set_samples_per_pass(self._h, n), self._h)

    def set_option(self, key, value):
        check(lib().rd_set_option(self._h, key.encode(), int(value)), self._h)

    def get_option(self, key):
        v = C.c_int64()
        check(lib().rd_get_option(self._h, key.encode(), C.byref(v)), self._h)
        return v.value

    def accumulator(self):
        sc = self._scene
        out = np.zeros((sc.height * sc.width, 4), np.float32)
        check(lib().rd_read_accumulator(self._h, _ptr(out), out.size), self._h)
        return out

    def counters(self):
        c = abi.RdCounters()
        check(lib().rd_get_counters(self._h, C.byref(c)), self._h)
        return {n: getattr(c, n) for n, _ in abi.RdCounters._fields_}

    def reset_counters(self):
        check(lib().rd_reset_counters(self._h), self._h)

    def kernel_times(self):
        t = abi.RdKernelTimes()
        check(lib().rd_get_kernel_times(self._h, C.byref(t)), self._h)
        return {n: (t.total_ms[i], t.launches[i]) for i, n in enumerate(abi.KERNEL_NAMES)}


# ------------------------------------------------------------- camera ------
def camera_update_matrices(camera):
    """Camera::update_matrices (camera.rs:210-231) through the C ABI: a copy
    of the camera dict (Scene.camera layout) with the four matrices
    recomputed in FP32."""
    s = Scene(camera, None, np.zeros(0, abi.OBJECT_DTYPE)).as_rd_camera()
    check(lib().rd_camera_update_matrices(C.byref(s)))
    out = dict(camera)
    for k in ("view_matrix", "proj_matrix", "inverse_view_matrix", "inverse_proj_matrix"):
        out[k] = list(getattr(s, k))
    return out


def camera_vulkan_inverse_view(camera):
    """The Vulkan renderer's inverse view (vulkan.rs:251-272) through the C
    ABI: 16 floats, column-major."""
    s = Scene(camera, None, np.zeros(0, abi.OBJECT_DTYPE)).as_rd_camera()
    out = np.zeros(16, np.float32)
    check(lib().rd_camera_vulkan_inverse_view(C.byref(s), out.ctypes.data_as(C.c_void_p)))
    return out


# -------------------------------------------------------------- .rscn ------
def load_rscn_native(path):
    """Parse a .rscn file with the library's native reader."""
    L = lib()
    p = C.POINTER(abi.RdScene)()
    check(L.rd_scene_load_rscn(path.encode(), C.byref(p)))
    try:
        return Scene.from_rd(p.contents)
    finally:
        L.rd_scene_free(p)


def save_rscn_native(scene, path):
    s = scene.as_rd()
    check(lib().rd_scene_save_rscn(C.byref(s), path.encode()))


# ------------------------------------------------------- parity hooks ------
def debug_trace(scene, rays, leaf=4, brute=False):
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    n = len(rays)
    ids = np.zeros(n, np.int32)
    ts = np.zeros(n, np.float32)
    normals = np.zeros((n, 3), np.float32)
    front = np.zeros(n, np.uint8)
    s = scene.as_rd()
    check(lib().rd_debug_trace(C.byref(s), leaf, int(brute), _ptr(rays), n, _ptr(ids), _ptr(ts),
                               _ptr(normals), _ptr(front)))
    return ids, ts, normals, front


def debug_extend(scene, rays, leaf=1, bvh4=1, lds=0, wide=0, jobs=1, refill=48, block=256, grid=0,
                 lane_chunk=0):
    """Closest hit through the production extend kernel (see rd_debug_extend).
    Returns (ids, ts, normals, front), or None when the renderer never launches
    that kernel form for this scene (RD_ERR_UNSUPPORTED)."""
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    n = len(rays)
    ids = np.zeros(n, np.int32)
    ts = np.zeros(n, np.float32)
    normals = np.zeros((n, 3), np.float32)
    front = np.zeros(n, np.uint8)
    s = scene.as_rd()
    v = abi.RdExtendVariant(bvh4, lds, wide, jobs, refill, block, grid, lane_chunk)
    st = lib().rd_debug_extend(C.byref(s), leaf, C.byref(v), _ptr(rays), n, _ptr(ids), _ptr(ts),
                               _ptr(normals), _ptr(front))
    if st == abi.RD_ERR_UNSUPPORTED:
        return None
    check(st)
    return ids, ts, normals, front


def debug_scatter(scene, packed, bits):
    packed = np.ascontiguousarray(packed, np.float32).reshape(-1, 16)
    bits = np.ascontiguousarray(bits, np.uint32).reshape(-1, DRAW_WORDS)
    out = np.zeros_like(packed)
    s = scene.as_rd()
    check(lib().rd_debug_scatter(C.byref(s), _ptr(packed), _ptr(bits), len(packed), _ptr(out)))
    return out


def debug_camera_rays(scene, pixels):
    pixels = np.ascontiguousarray(pixels, np.uint32)
    out = np.zeros((len(pixels), 6), np.float32)
    s = scene.as_rd()
    check(lib().rd_debug_camera_rays(C.byref(s), _ptr(pixels), len(pixels), _ptr(out)))
    return out


def debug_world(scene, dirs):
    dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    out = np.zeros_like(dirs)
    s = scene.as_rd()
    check(lib().rd_debug_world(C.byref(s), _ptr(dirs), len(dirs), _ptr(out)))
    return out


def debug_copy_calibration(nbytes, reps=5):
    """Runs the library's plain copy kernel `reps` times over nbytes; returns
    each launch's duration in ms."""
    ms = np.zeros(reps, np.float32)
    check(lib().rd_debug_copy_calibration(int(nbytes), int(reps), _ptr(ms)))
    return ms


def debug_philox(ctr_key):
    ctr_key = np.ascontiguousarray(ctr_key, np.uint32).reshape(-1, 6)
    out = np.zeros((len(ctr_key), 4), np.uint32)
    check(lib().rd_debug_philox(_ptr(ctr_key), len(ctr_key), _ptr(out)))
    return out


def host_bvh_info(scene, leaf=4):
    n = C.c_uint32()
    d = C.c_uint32()
    err = C.create_string_buffer(256)
    s = scene.as_rd()
    st = lib().rd_debug_host_bvh_info(C.byref(s), leaf, C.byref(n), C.byref(d), err, 256)
    return st, n.value, d.value, err.value.decode()


def ring_ops(capacity, ops):
    """Replays reserve (>= 0) / release (< 0) operations on the radiance ring."""
    ops = np.ascontiguousarray(ops, np.int64)
    out = np.zeros(len(ops), np.int64)
    check(lib().rd_debug_ring_ops(capacity, _ptr(ops), len(ops), _ptr(out)))
    return out


def cache_classes(nbytes):
    """Size classes of the reused-buffer cache for requests of nbytes."""
    nbytes = np.ascontiguousarray(nbytes, np.uint64)
    out = np.zeros(len(nbytes), np.uint64)
    check(lib().rd_debug_cache_classes(_ptr(nbytes), len(nbytes), _ptr(out)))
    return out


def host_bvh4_info(scene, leaf=4):
    """(status, 4-wide node count, traversal stack bound, error text)."""
    n = C.c_uint32()
    d = C.c_uint32()
    err = C.create_string_buffer(256)
    s = scene.as_rd()
    st = lib().rd_debug_host_bvh4_info(C.byref(s), leaf, C.byref(n), C.byref(d), err, 256)
    return st, n.value, d.value, err.value.decode()


def host_trace(scene, rays, leaf=4, width=2):
    """Closest hit through the binary tree (width 2) or the 4-wide tree."""
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    ids = np.zeros(len(rays), np.int32)
    ts = np.zeros(len(rays), np.float32)
    s = scene.as_rd()
    fn = {2: lib().rd_debug_host_trace, 4: lib().rd_debug_host_trace4}[width]
    check(fn(C.byref(s), leaf, _ptr(rays), len(rays), _ptr(ids), _ptr(ts)))
    return ids, ts


def host_scatter(scene, packed, bits):
    """The kernels' scatter compiled for the host (same layout as debug_scatter)."""
    packed = np.ascontiguousarray(packed, np.float32).reshape(-1, 16)
    bits = np.ascontiguousarray(bits, np.uint32).reshape(-1, DRAW_WORDS)
    out = np.zeros_like(packed)
    s = scene.as_rd()
    check(lib().rd_debug_host_scatter(C.byref(s), _ptr(packed), _ptr(bits), len(packed), _ptr(out)))
    return out


def host_camera_rays(scene, pixels):
    pixels = np.ascontiguousarray(pixels, np.uint32)
    out = np.zeros((len(pixels), 6), np.float32)
    s = scene.as_rd()
    check(lib().rd_debug_host_camera_rays(C.byref(s), _ptr(pixels), len(pixels), _ptr(out)))
    return out
