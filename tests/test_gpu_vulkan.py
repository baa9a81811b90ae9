"""The Vulkan-semantics render mode (option "semantics" = 1) through the C
ABI against the oracle's restatement of the reference's Vulkan renderer
(shaders/vulkan/raytrace.rgen and friends, src/renderer/vulkan.rs): the
accumulators bit for bit and the rounded UNORM images byte for byte, over
the closest-hit kernel's tree layouts, plus the mode's render_sample and
option rules.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
import raydar_amd as ra
from raydar_amd import abi, scenegen

pytestmark = pytest.mark.gpu


def _vk_oracle(sc, spp, bounces, pixels=None):
    iv = ra.camera_vulkan_inverse_view(sc.camera)
    return oracle.vk_render(sc, iv, spp, bounces, pixels=pixels, threads=8)


def _gpu(sc, spp, bounces, **opts):
    with ra.HipRenderer(max_sample_count=spp, max_bounces=bounces, seed=9, semantics=1, **opts) as r:
        img = r.render_frame(sc)
        acc = r.accumulator()
        assert r.sample_count == spp
    return img, acc


def _check(sc, spp, bounces, **opts):
    print(f"vk frame {sc.width}x{sc.height} objects={len(sc.objects)} spp={spp} opts={opts}", flush=True)
    img, acc = _gpu(sc, spp, bounces, **opts)
    ref = _vk_oracle(sc, spp, bounces)
    np.testing.assert_array_equal(acc.view(np.uint32), ref.view(np.uint32))
    np.testing.assert_array_equal(img.reshape(-1), oracle.vk_tonemap(ref, spp))


@pytest.mark.parametrize("opts", [dict(), dict(bvh4=0), dict(lds_bvh=0), dict(bvh4=0, lds_bvh=0),
                                  dict(traversal=1), dict(leaf=4), dict(refill=16, lds_objects=0),
                                  dict(queue=4096, grid=5)])
def test_vulkan_frame_matches_oracle(opts):
    _check(scenegen.plain_small(0), 8, 12, **opts)


def test_vulkan_material_zoo_matches_oracle():
    # every scatter branch: glass (refracting and reflecting), metal, diffuse,
    # specular, emitters; cubes and spheres
    from tests import raygen
    _check(raygen.material_zoo(), 6, 12)


@pytest.mark.parametrize("bvh4", [0, 1])
def test_vulkan_clustered_scene_matches_oracle(bvh4):
    # 4096 objects: the tree stays in global memory
    _check(scenegen.clustered(2, 4096, 96, 54), 2, 12, bvh4=bvh4)


def test_vulkan_render_sample_renders_every_sample_at_once():
    sc = scenegen.plain_small(1)
    with ra.HipRenderer(max_sample_count=5, max_bounces=8, semantics=1) as r:
        r.new_frame(sc)
        img = r.render_sample(sc)
        assert img is not None and r.sample_count == 5
        assert r.render_sample(sc) is None  # vulkan.rs:112-114
        acc = r.accumulator()
    ref = _vk_oracle(sc, 5, 8)
    np.testing.assert_array_equal(acc.view(np.uint32), ref.view(np.uint32))
    np.testing.assert_array_equal(img.reshape(-1), oracle.vk_tonemap(ref, 5))


def test_vulkan_and_cpu_semantics_alternate_on_one_renderer():
    sc = scenegen.plain_small(2)
    imgs = []
    with ra.HipRenderer(max_sample_count=4, max_bounces=12, seed=3) as r:
        for vk in (1, 0, 1):
            r.set_option("semantics", vk)
            img = r.render_frame(sc)
            acc = r.accumulator()
            if vk:
                ref = _vk_oracle(sc, 4, 12)
                want = oracle.vk_tonemap(ref, 4)
            else:
                ref = oracle.render(sc, 4, 12, 3)
                want = oracle.tonemap(ref, 4)
            np.testing.assert_array_equal(acc.view(np.uint32), ref.view(np.uint32))
            np.testing.assert_array_equal(img.reshape(-1), want)
            imgs.append(img.copy())
    # the two semantics give different images; the same one twice, the same image
    assert np.mean(imgs[0] != imgs[1]) > 0.5
    np.testing.assert_array_equal(imgs[0], imgs[2])


def test_vulkan_mode_option_rules():
    sc = scenegen.plain_small(0)
    with ra.HipRenderer(max_sample_count=2, max_bounces=4) as r:
        assert r.get_option("semantics") == 0
        with pytest.raises(ra.RaydarError):
            r.set_option("semantics", 2)
        # changed between new_frame and the samples: refused, then fine after a new frame
        r.new_frame(sc)
        r.set_option("semantics", 1)
        with pytest.raises(ra.RaydarError, match="semantics"):
            r.render_sample(sc)
        r.new_frame(sc)
        assert r.render_sample(sc) is not None
        # launch shapes the mode has no kernels for
        r.set_option("block", 128)
        with pytest.raises(ra.RaydarError, match="block 256"):
            r.render_frame(sc)
        r.set_option("block", 256)
        r.set_option("count", 1)
        with pytest.raises(ra.RaydarError, match="counting"):
            r.render_frame(sc)
        r.set_option("count", 0)
        # the allocation-free entry point keeps to the CPU renderer's semantics
        L = ra.lib()
        L.rd_render_frame_noalloc.restype = C.c_int
        L.rd_render_frame_noalloc.argtypes = [C.c_void_p, C.POINTER(abi.RdScene), C.c_void_p,
                                              C.c_size_t]
        out = np.zeros((sc.height, sc.width, 4), np.uint8)
        s = sc.as_rd()
        assert L.rd_render_frame_noalloc(r._h, C.byref(s), out.ctypes.data_as(C.c_void_p),
                                         out.nbytes) == abi.RD_ERR_UNSUPPORTED
        # and the renderer still renders
        img = r.render_frame(sc)
    ref = _vk_oracle(sc, 2, 4)
    np.testing.assert_array_equal(img.reshape(-1), oracle.vk_tonemap(ref, 2))
