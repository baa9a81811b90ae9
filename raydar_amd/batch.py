"""Python access to the batched renderer and the render server
(include/raydar_hip_batch.h) for tests and benchmarks."""
import ctypes as C

import numpy as np

from . import abi, check, lib


class RdJobDesc(C.Structure):
    _fields_ = [("scene", C.POINTER(abi.RdScene)), ("samples", C.c_uint32),
                ("max_bounces", C.c_uint32), ("seed", C.c_uint64), ("pass_samples", C.c_uint32),
                ("rgba", C.c_void_p), ("accumulator", C.c_void_p), ("progress_rgba", C.c_void_p)]


class RdBatchStats(C.Structure):
    _fields_ = [("elapsed_ns", C.c_int64), ("paths", C.c_uint64), ("iterations", C.c_uint64),
                ("graph_launches", C.c_uint64), ("pumps", C.c_uint64), ("build_ns", C.c_int64),
                ("host_threads", C.c_uint32)]


class RdTuneResult(C.Structure):
    _fields_ = [("block", C.c_int32), ("leaf", C.c_int32), ("queue", C.c_uint32),
                ("best_ms", C.c_double), ("default_ms", C.c_double),
                ("configs_tried", C.c_uint32), ("cached", C.c_uint32)]


class RdServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("pass_samples", C.c_uint32),
                ("block", C.c_int32), ("queue", C.c_uint32), ("leaf", C.c_int32)]


class RdJobResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("done", C.c_int32), ("samples_done", C.c_uint32),
                ("passes_done", C.c_uint32), ("queued_ns", C.c_int64), ("latency_ns", C.c_int64),
                ("device_bytes", C.c_uint64)]


class RdServerStats(C.Structure):
    _fields_ = [("jobs_submitted", C.c_uint64), ("jobs_done", C.c_uint64),
                ("jobs_rejected", C.c_uint64), ("paths", C.c_uint64), ("iterations", C.c_uint64),
                ("busy_ns", C.c_int64), ("hbm_in_use", C.c_uint64), ("hbm_peak", C.c_uint64),
                ("engine_bytes", C.c_uint64), ("jobs_resident", C.c_uint32), ("jobs_queued", C.c_uint32),
                ("max_resident", C.c_uint32)]


_bound = False


def _bind():
    global _bound
    if _bound:
        return lib()
    L = lib()
    R = C.c_void_p
    P = C.POINTER
    sig = {
        "rd_batch_create": (C.c_int, [P(R)]),
        "rd_batch_destroy": (None, [R]),
        "rd_batch_set_option": (C.c_int, [R, C.c_char_p, C.c_int64]),
        "rd_batch_get_option": (C.c_int, [R, C.c_char_p, P(C.c_int64)]),
        "rd_batch_render": (C.c_int, [R, P(RdJobDesc), C.c_uint32]),
        "rd_batch_stats_get": (C.c_int, [R, P(RdBatchStats)]),
        "rd_batch_last_error": (C.c_char_p, [R]),
        "rd_batch_autotune": (C.c_int, [R, P(RdJobDesc), C.c_uint32, P(RdTuneResult)]),
        "rd_server_create": (C.c_int, [P(RdServerConfig), P(R)]),
        "rd_server_destroy": (None, [R]),
        "rd_server_submit": (C.c_int, [R, P(RdJobDesc), P(C.c_uint64)]),
        "rd_server_poll": (C.c_int, [R, C.c_uint64, P(RdJobResult)]),
        "rd_server_wait": (C.c_int, [R, C.c_uint64, C.c_int64, P(RdJobResult)]),
        "rd_server_stats_get": (C.c_int, [R, P(RdServerStats)]),
        "rd_server_last_error": (C.c_char_p, [R]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _bound = True
    return L


class Job:
    """A render job plus its output buffers (kept alive with the job)."""

    def __init__(self, scene, samples, max_bounces=12, seed=0, pass_samples=0, progress=False,
                 accumulator=True):
        self.scene = scene
        self.rd_scene = scene.as_rd()
        self.rgba = np.zeros((scene.height, scene.width, 4), np.uint8)
        self.acc = np.zeros((scene.height * scene.width, 4), np.float32) if accumulator else None
        self.progress = np.zeros_like(self.rgba) if progress else None
        self.desc = RdJobDesc(C.pointer(self.rd_scene), samples, max_bounces, seed, pass_samples,
                              self.rgba.ctypes.data,
                              self.acc.ctypes.data if accumulator else None,
                              self.progress.ctypes.data if progress else None)

    @property
    def paths(self):
        return self.scene.width * self.scene.height * self.desc.samples


def _err(L, fn, h):
    m = fn(h)
    return m.decode() if m else ""


class BatchRenderer:
    def __init__(self, **options):
        L = _bind()
        h = C.c_void_p()
        check(L.rd_batch_create(C.byref(h)))
        self._h = h
        for k, v in options.items():
            self.set_option(k, v)

    def _check(self, st):
        if st != abi.RD_OK:
            from . import RaydarError
            raise RaydarError(st, _err(lib(), lib().rd_batch_last_error, self._h))

    def set_option(self, key, value):
        self._check(lib().rd_batch_set_option(self._h, key.encode(), int(value)))

    def get_option(self, key):
        v = C.c_int64()
        self._check(lib().rd_batch_get_option(self._h, key.encode(), C.byref(v)))
        return v.value

    def render(self, jobs):
        arr = (RdJobDesc * len(jobs))(*[j.desc for j in jobs])
        self._check(lib().rd_batch_render(self._h, arr, len(jobs)))
        return self.stats()

    def stats(self):
        s = RdBatchStats()
        self._check(lib().rd_batch_stats_get(self._h, C.byref(s)))
        return {n: getattr(s, n) for n, _ in RdBatchStats._fields_}

    def autotune(self, jobs):
        arr = (RdJobDesc * len(jobs))(*[j.desc for j in jobs])
        r = RdTuneResult()
        self._check(lib().rd_batch_autotune(self._h, arr, len(jobs), C.byref(r)))
        return {n: getattr(r, n) for n, _ in RdTuneResult._fields_}

    def close(self):
        if self._h:
            lib().rd_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RenderServer:
    def __init__(self, hbm_budget_bytes=200 << 30, pass_samples=8, block=0, queue=0, leaf=0):
        L = _bind()
        cfg = RdServerConfig(hbm_budget_bytes, pass_samples, block, queue, leaf)
        h = C.c_void_p()
        check(L.rd_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._jobs = {}

    def _check(self, st):
        if st != abi.RD_OK:
            from . import RaydarError
            raise RaydarError(st, _err(lib(), lib().rd_server_last_error, self._h))

    def submit(self, job):
        jid = C.c_uint64()
        self._check(lib().rd_server_submit(self._h, C.byref(job.desc), C.byref(jid)))
        self._jobs[jid.value] = job
        return jid.value

    def _res(self, r):
        return {n: getattr(r, n) for n, _ in RdJobResult._fields_}

    def poll(self, jid):
        r = RdJobResult()
        self._check(lib().rd_server_poll(self._h, jid, C.byref(r)))
        return self._res(r)

    def wait(self, jid, timeout_ms=-1):
        r = RdJobResult()
        self._check(lib().rd_server_wait(self._h, jid, timeout_ms, C.byref(r)))
        return self._res(r)

    def stats(self):
        s = RdServerStats()
        self._check(lib().rd_server_stats_get(self._h, C.byref(s)))
        return {n: getattr(s, n) for n, _ in RdServerStats._fields_}

    def close(self):
        if self._h:
            lib().rd_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
