"""ctypes access to the oracle library (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it as the checker or the timed CPU baseline.  The
product package raydar_amd never imports it.
"""
import ctypes as C
import os

import numpy as np

from raydar_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

_P = np.ctypeslib.ndpointer


def build():
    """Compile liboracle.so / liboracle_fast.so with g++ (see Makefile)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "all"], check=True)


def lib(fast=False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name in _LIBS:
        return _LIBS[name]
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        build()
    _LIBS[name] = _bind(C.CDLL(path))
    return _LIBS[name]


def build_mutant(macro, out_dir):
    """The oracle compiled with one planted mistake (an ORACLE_MUTANT_* macro
    of oracle.cpp) into out_dir: the tests use it to show that a pin fails
    when the oracle makes that mistake."""
    import subprocess
    path = os.path.join(str(out_dir), f"liboracle_{macro.lower()}.so")
    subprocess.run(["g++", "-O1", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-std=c++17",
                    "-shared", f"-D{macro}", "-o", path, os.path.join(_HERE, "oracle.cpp")],
                   check=True)
    return _bind(C.CDLL(path))


def _bind(L):
    sp = C.POINTER(abi.RdScene)
    L.or_philox4x32_10.argtypes = [_P(np.uint32), _P(np.uint32), _P(np.uint32)]
    L.or_bounce_bits.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, _P(np.uint32)]
    L.or_camera_rays_f32.argtypes = [sp, _P(np.uint32), C.c_uint32, _P(np.float32)]
    L.or_trace_f32.argtypes = [sp, _P(np.float32), C.c_uint32, _P(np.int32), _P(np.float32),
                               _P(np.float32), _P(np.uint8)]
    L.or_scatter_f32.argtypes = [sp, _P(np.float32), _P(np.uint32), C.c_uint32, _P(np.float32)]
    L.or_world_f32.argtypes = [sp, _P(np.float32), C.c_uint32, _P(np.float32)]
    L.or_render_f32.argtypes = [sp, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, C.c_void_p,
                                C.c_uint32, _P(np.float32), C.c_int]
    L.or_render_f64.argtypes = [sp, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, C.c_void_p,
                                C.c_uint32, _P(np.float64), C.c_int]
    L.or_tonemap.argtypes = [_P(np.float32), C.c_uint64, C.c_uint32, _P(np.uint8)]
    L.or_vk_pcg.argtypes = [C.c_uint32, C.c_uint32, _P(np.uint32)]
    L.or_vk_camera_rays.argtypes = [sp, _P(np.float32), _P(np.uint32), C.c_uint32, _P(np.float32)]
    L.or_vk_trace.argtypes = [sp, _P(np.float32), C.c_uint32, _P(np.int32), _P(np.float32),
                              _P(np.float32), _P(np.uint8)]
    L.or_vk_scatter.argtypes = [sp, _P(np.float32), _P(np.float32), _P(np.float32)]
    L.or_vk_world.argtypes = [sp, _P(np.float32), C.c_uint32, _P(np.float32)]
    L.or_vk_render.argtypes = [sp, _P(np.float32), C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32,
                               _P(np.float32), C.c_int]
    L.or_vk_tonemap.argtypes = [_P(np.float32), C.c_uint64, C.c_uint32, _P(np.uint8)]
    for f in ("or_philox4x32_10", "or_bounce_bits", "or_camera_rays_f32", "or_trace_f32",
              "or_scatter_f32", "or_world_f32", "or_render_f32", "or_render_f64", "or_tonemap",
              "or_vk_pcg", "or_vk_camera_rays", "or_vk_trace", "or_vk_scatter", "or_vk_world",
              "or_vk_render", "or_vk_tonemap"):
        getattr(L, f).restype = None
    return L


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(_c(ctr, np.uint32), _c(key, np.uint32), out)
    return out


DRAW_WORDS = 8  # random words per bounce: two Philox-4x32 blocks


def bounce_bits(seed, pixel, sample, bounce):
    out = np.zeros(DRAW_WORDS, np.uint32)
    lib().or_bounce_bits(seed, pixel, sample, bounce, out)
    return out


def camera_rays(scene, pixels):
    pixels = _c(pixels, np.uint32)
    out = np.zeros((len(pixels), 6), np.float32)
    s = scene.as_rd()
    lib().or_camera_rays_f32(C.byref(s), pixels, len(pixels), out)
    return out


def trace(scene, rays):
    rays = _c(rays, np.float32).reshape(-1, 6)
    n = len(rays)
    ids = np.zeros(n, np.int32)
    ts = np.zeros(n, np.float32)
    normals = np.zeros((n, 3), np.float32)
    front = np.zeros(n, np.uint8)
    s = scene.as_rd()
    lib().or_trace_f32(C.byref(s), rays, n, ids, ts, normals, front)
    return ids, ts, normals, front


def scatter(scene, packed, bits, L=None):
    """packed: n x 16 floats (see oracle.cpp or_scatter_f32); bits: n x 8.
    L: another build of the oracle (build_mutant), default the shipped one."""
    packed = _c(packed, np.float32).reshape(-1, 16)
    bits = _c(bits, np.uint32).reshape(-1, DRAW_WORDS)
    out = np.zeros_like(packed)
    s = scene.as_rd()
    (L or lib()).or_scatter_f32(C.byref(s), packed, bits, len(packed), out)
    return out


def world(scene, dirs):
    dirs = _c(dirs, np.float32).reshape(-1, 3)
    out = np.zeros_like(dirs)
    s = scene.as_rd()
    lib().or_world_f32(C.byref(s), dirs, len(dirs), out)
    return out


def render(scene, spp, max_bounces, seed, sample_base=0, pixels=None, fp64=False, threads=None,
           fast=False):
    """Accumulator (n x 4) of `spp` samples for `pixels` (all when None)."""
    if threads is None:
        threads = os.cpu_count() or 1
    n = scene.width * scene.height if pixels is None else len(pixels)
    acc = np.zeros((n, 4), np.float64 if fp64 else np.float32)
    pix = None
    if pixels is not None:
        pix = _c(pixels, np.uint32)
    s = scene.as_rd()
    L = lib(fast)
    fn = L.or_render_f64 if fp64 else L.or_render_f32
    fn(C.byref(s), spp, max_bounces, seed, sample_base,
       None if pix is None else pix.ctypes.data_as(C.c_void_p), 0 if pix is None else len(pix),
       acc, threads)
    return acc


def tonemap(acc, sample_count):
    acc = _c(acc, np.float32)
    out = np.zeros(acc.size, np.uint8)
    lib().or_tonemap(acc.reshape(-1), acc.size // 4, sample_count, out)
    return out


# ------------------------------------------------ Vulkan-renderer semantics ---
# vk_inv_view: the inverse of look_at_rh(position, target, -up) (vulkan.rs:252-272),
# 16 floats column-major; the tests take it from raydar_amd.camera_vulkan_inverse_view.

def vk_pcg(state, n):
    out = np.zeros(n, np.uint32)
    lib().or_vk_pcg(state, n, out)
    return out


def vk_camera_rays(scene, vk_inv_view, pixels):
    pixels = _c(pixels, np.uint32)
    out = np.zeros((len(pixels), 6), np.float32)
    s = scene.as_rd()
    lib().or_vk_camera_rays(C.byref(s), _c(vk_inv_view, np.float32), pixels, len(pixels), out)
    return out


def vk_trace(scene, rays):
    rays = _c(rays, np.float32).reshape(-1, 6)
    n = len(rays)
    ids = np.zeros(n, np.int32)
    ts = np.zeros(n, np.float32)
    normals = np.zeros((n, 3), np.float32)
    front = np.zeros(n, np.uint8)
    s = scene.as_rd()
    lib().or_vk_trace(C.byref(s), rays, n, ids, ts, normals, front)
    return ids, ts, normals, front


def vk_scatter(scene, packed, draws):
    """packed: 14 floats (ray o, d, t, object id bits, atten, light); draws:
    the uniforms, in order.  Returns 14 floats: next o, d, atten, light,
    transmission flag, number of uniforms used."""
    packed = _c(packed, np.float32).reshape(-1)
    d = np.zeros(16, np.float32)
    d[:len(draws)] = draws
    out = np.zeros(14, np.float32)
    s = scene.as_rd()
    lib().or_vk_scatter(C.byref(s), packed, d, out)
    return out


def vk_world(scene, dirs):
    dirs = _c(dirs, np.float32).reshape(-1, 3)
    out = np.zeros_like(dirs)
    s = scene.as_rd()
    lib().or_vk_world(C.byref(s), dirs, len(dirs), out)
    return out


def vk_render(scene, vk_inv_view, spp, max_bounces, pixels=None, threads=None):
    """Accumulator (n x 4) of all `spp` samples of `pixels` (all when None)."""
    if threads is None:
        threads = os.cpu_count() or 1
    n = scene.width * scene.height if pixels is None else len(pixels)
    acc = np.zeros((n, 4), np.float32)
    pix = None if pixels is None else _c(pixels, np.uint32)
    s = scene.as_rd()
    lib().or_vk_render(C.byref(s), _c(vk_inv_view, np.float32), spp, max_bounces,
                       None if pix is None else pix.ctypes.data_as(C.c_void_p),
                       0 if pix is None else len(pix), acc, threads)
    return acc


def vk_tonemap(acc, sample_count):
    acc = _c(acc, np.float32)
    out = np.zeros(acc.size, np.uint8)
    lib().or_vk_tonemap(acc.reshape(-1), acc.size // 4, sample_count, out)
    return out

