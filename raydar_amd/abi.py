"""ctypes mirror of include/raydar_hip.h.

The structs follow raydar's serde data model: rd_object = Object + Material
(src/scene/objects.rs:13-17, src/scene/material.rs:4-13), rd_camera = Camera
(src/scene/camera.rs:13-30), rd_world = World (src/scene/world.rs:6-14),
rd_scene = Scene (src/scene/mod.rs:13-18), rd_config = RendererConfig
(src/renderer/mod.rs:11-23), rd_timings = Profiler (src/renderer/timing.rs:9-19).
"""
import ctypes as C

import numpy as np

RD_OK = 0
RD_ERR_INVALID_ARGUMENT = 1
RD_ERR_UNSUPPORTED = 2
RD_ERR_OUT_OF_MEMORY = 3
RD_ERR_DEVICE = 4
RD_ERR_BUFFER_TOO_SMALL = 5
RD_ERR_PARSE = 6
RD_ERR_IO = 7
RD_ERR_BUDGET = 8
RD_ERR_NOT_FOUND = 9

STATUS_NAMES = {
    0: "RD_OK", 1: "RD_ERR_INVALID_ARGUMENT", 2: "RD_ERR_UNSUPPORTED",
    3: "RD_ERR_OUT_OF_MEMORY", 4: "RD_ERR_DEVICE", 5: "RD_ERR_BUFFER_TOO_SMALL",
    6: "RD_ERR_PARSE", 7: "RD_ERR_IO", 8: "RD_ERR_BUDGET", 9: "RD_ERR_NOT_FOUND",
}

RD_SPHERE = 0
RD_CUBE = 1
RD_WORLD_SKY_COLOR = 0
RD_WORLD_SOLID_COLOR = 1
RD_WORLD_TRANSPARENT = 2
RD_PROJECTION_PERSPECTIVE = 0
RD_PROJECTION_ORTHOGRAPHIC = 1

F3 = C.c_float * 3
F16 = C.c_float * 16


class RdObject(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32), ("center", F3), ("size", C.c_float), ("albedo", F3),
        ("roughness", C.c_float), ("metallic", C.c_float), ("emission_color", F3),
        ("emission_strength", C.c_float), ("transmission", C.c_float), ("ior", C.c_float),
    ]


# numpy view of an rd_object array (64 bytes per object)
OBJECT_DTYPE = np.dtype([
    ("kind", "<u4"), ("center", "<f4", 3), ("size", "<f4"), ("albedo", "<f4", 3),
    ("roughness", "<f4"), ("metallic", "<f4"), ("emission_color", "<f4", 3),
    ("emission_strength", "<f4"), ("transmission", "<f4"), ("ior", "<f4"),
])
assert OBJECT_DTYPE.itemsize == C.sizeof(RdObject) == 64


class RdCamera(C.Structure):
    _fields_ = [
        ("position", F3), ("target", F3), ("up", F3),
        ("resolution_x", C.c_uint32), ("resolution_y", C.c_uint32),
        ("projection", C.c_uint32), ("projection_param", C.c_float),
        ("near_clip", C.c_float), ("far_clip", C.c_float),
        ("view_matrix", F16), ("proj_matrix", F16),
        ("inverse_view_matrix", F16), ("inverse_proj_matrix", F16),
    ]


class RdWorld(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("top_color", F3), ("bottom_color", F3),
                ("solid_color", F3)]


class RdScene(C.Structure):
    _fields_ = [("camera", RdCamera), ("world", RdWorld), ("object_count", C.c_uint32),
                ("objects", C.POINTER(RdObject))]


class RdConfig(C.Structure):
    _fields_ = [("max_sample_count", C.c_uint32), ("max_bounces", C.c_uint32)]


class RdTimings(C.Structure):
    _fields_ = [("frame_ns", C.c_int64), ("sample_ns", C.c_int64),
                ("prepare_ns", C.c_int64), ("render_ns", C.c_int64)]


class RdCounters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "camera_rays", "segments", "box_tests", "sphere_tests", "cube_tests",
        "sphere_hits", "cube_hits", "glass_scatters", "misses", "accumulates",
        "tonemapped_px", "wave_steps", "lane_steps", "sort_steps", "prim_trips", "cube_trips",
        "refills", "reused_hits")]


class RdKernelTimes(C.Structure):
    _fields_ = [("total_ms", C.c_double * 5), ("launches", C.c_uint64 * 5)]


class RdExtendVariant(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("bvh4", "lds", "wide", "jobs", "refill", "block", "grid",
                                         "lane_chunk")]


KERNEL_NAMES = ("generate", "extend", "shade", "accumulate", "tonemap")

# Every symbol the C ABI exports (include/raydar_hip.h, include/raydar_hip_batch.h,
# include/raydar_hip_debug.h).  Checked by the CPU test-suite.
EXPORTED_SYMBOLS = (
    "rd_config_default", "rd_create", "rd_destroy", "rd_new_frame", "rd_render_sample",
    "rd_render_frame", "rd_profiler", "rd_sample_count", "rd_max_sample_count",
    "rd_max_bounces", "rd_set_max_sample_count", "rd_set_max_bounces", "rd_last_error",
    "rd_set_samples_per_pass", "rd_set_seed", "rd_read_accumulator", "rd_restart_frame",
    "rd_set_option",
    "rd_get_option", "rd_get_counters", "rd_reset_counters", "rd_get_kernel_times",
    "rd_scene_load_rscn", "rd_scene_parse_rscn", "rd_scene_save_rscn", "rd_scene_free",
)
