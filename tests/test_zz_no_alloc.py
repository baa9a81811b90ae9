"""rd_render_frame_noalloc (the allocation-free variant of rd_render_frame,
whose device buffers all come from one static array per GPU) against
rd_render_frame through the C ABI: identical images, sample count and
timers on two calls in a row, the second with a smaller scene; no change in
the device's free memory across calls; a scene beyond the array's limit is
refused with a clear error and leaves the renderer usable."""
import ctypes as C

import numpy as np
import pytest

import raydar_amd as ra
from raydar_amd import abi, scenegen

pytestmark = pytest.mark.gpu


def _noalloc(r, scene):
    out = np.zeros((scene.height, scene.width, 4), np.uint8)
    s = scene.as_rd()
    st = ra.lib().rd_render_frame_noalloc(r._h, C.byref(s), out.ctypes.data_as(C.c_void_p), out.nbytes)
    return st, out


def _device_free_bytes():
    hip = C.CDLL("libamdhip64.so.7")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_noalloc_frame_equals_render_frame():
    ra.lib().rd_render_frame_noalloc.restype = C.c_int
    ra.lib().rd_render_frame_noalloc.argtypes = [C.c_void_p, C.POINTER(abi.RdScene), C.c_void_p,
                                                 C.c_size_t]
    # 1920x1080 with 97 objects (3 passes of 2 samples from the leftover
    # radiance slots), then a smaller 256x256 scene
    cases = [(scenegen.plain_1080p(1), 6, 12, 7), (scenegen.plain_small(0), 16, 12, 5)]
    refs = []
    for sc, spp, b, seed in cases:
        with ra.HipRenderer(max_sample_count=spp, max_bounces=b, seed=seed) as ref:
            refs.append(ref.render_frame(sc))
    with ra.HipRenderer() as r:
        for (sc, spp, b, seed), want in zip(cases, refs):
            r.max_sample_count, r.max_bounces = spp, b
            r.set_seed(seed)
            st, out = _noalloc(r, sc)
            assert st == abi.RD_OK, ra.last_error(r._h)
            np.testing.assert_array_equal(out, want)
            assert r.sample_count == spp
            t = r.profiler()
            assert all(v is not None and v > 0 for v in t.values()), t
        # no device memory is taken by further calls
        free0 = _device_free_bytes()
        for (sc, spp, b, seed), want in zip(cases[::-1], refs[::-1]):
            r.max_sample_count, r.max_bounces = spp, b
            r.set_seed(seed)
            st, out = _noalloc(r, sc)
            assert st == abi.RD_OK
            np.testing.assert_array_equal(out, want)
        assert _device_free_bytes() >= free0 - (1 << 20)
        # beyond the limit: refused with the byte count, the renderer still works
        huge = scenegen.plain_small(0)
        huge.camera = dict(huge.camera, resolution_x=4096, resolution_y=4096)
        st, _ = _noalloc(r, huge)
        assert st == abi.RD_ERR_INVALID_ARGUMENT
        assert "static scratch" in ra.last_error(r._h)
        sc, spp, b, seed = cases[1]
        r.max_sample_count, r.max_bounces = spp, b
        r.set_seed(seed)
        st, out = _noalloc(r, sc)
        assert st == abi.RD_OK
        np.testing.assert_array_equal(out, refs[1])
