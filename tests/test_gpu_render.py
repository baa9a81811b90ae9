"""Whole-frame parity of the HIP renderer through the C ABI.

Images are compared with the oracle's FP32 render on the same draws (the
kernels reproduce its arithmetic, so accumulators agree to 1e-5 relative and
normally bit for bit), against closed forms, and, for the image-error check
of the path's contract, against an FP64 high-spp render: the GPU image's
per-pixel RMSE to that truth must be within 1% of the CPU FP32 renderer's at
equal spp (tests/golden/plain_small_truth.npz, made by make_truth.py).
"""
import os
import time

import numpy as np
import pytest

import oracle
import raydar_amd as ra
from raydar_amd import abi, scenegen
from raydar_amd.scene import Scene, make_camera, make_object, sky_world, solid_world

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _close(a, b, rel=1e-5):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) <= rel * np.maximum(np.maximum(np.abs(a), np.abs(b)), 1.0)


def _gpu_acc(scene, spp, bounces, seed, **opts):
    with ra.HipRenderer(max_sample_count=spp, max_bounces=bounces, seed=seed, **opts) as r:
        img = r.render_frame(scene)
        acc = r.accumulator()
        assert r.sample_count == spp
    return img, acc


@pytest.mark.parametrize("opts", [dict(), dict(leaf=1, block=64, queue=4096),
                                  dict(leaf=8, block=128, queue=65536), dict(traversal=1),
                                  dict(refill=0), dict(refill=1, grid=7), dict(refill=64, leaf=2),
                                  dict(refill=0, block=64, queue=4096), dict(lds_bvh=0),
                                  dict(lds_bvh=0, refill=32, block=128), dict(bvh4=0),
                                  dict(bvh4=0, lds_bvh=0), dict(bvh4=0, leaf=8, refill=16),
                                  dict(pipes=1), dict(pipes=3, queue=65536), dict(pipes=4, queue=4096),
                                  dict(pipes=4, sync_every=1, lds_bvh=0, queue=1 << 16),
                                  dict(lds_objects=0), dict(lds_objects=0, lds_bvh=0, pipes=2),
                                  dict(lane_chunk=0), dict(lane_chunk=64), dict(lane_chunk=64, refill=1, grid=7),
                                  dict(lane_chunk=4096, lds_bvh=0, block=128)])
def test_plain_small_matches_oracle(opts):
    sc = scenegen.plain_small(0)
    img, acc = _gpu_acc(sc, 16, 12, 5, **opts)
    ref = oracle.render(sc, 16, 12, 5)
    assert np.all(_close(acc, ref))
    exact = np.mean(acc.view(np.uint32) == ref.view(np.uint32))
    print(f"opts={opts}: accumulator bit-exact fraction {exact:.6f}")
    assert exact == 1.0
    np.testing.assert_array_equal(img.reshape(-1), oracle.tonemap(ref, 16))


@pytest.mark.parametrize("size", [(37, 29), (5, 3), (61, 8), (16, 17)])
def test_image_sizes_off_the_tile_grid_match_oracle(size):
    # camera paths are generated in 8x8 pixel tiles; a width that is not a
    # multiple of 8 leaves a column strip and a height a last band of fewer
    # than 8 rows, each taken row-major: every pixel must still get exactly
    # its own samples
    sc = scenegen.plain_1080p(1, *size)
    print(f"frame {size[0]}x{size[1]}", flush=True)
    img, acc = _gpu_acc(sc, 3, 12, 7)
    ref = oracle.render(sc, 3, 12, 7)
    np.testing.assert_array_equal(acc.view(np.uint32), ref.view(np.uint32))
    np.testing.assert_array_equal(img.reshape(-1), oracle.tonemap(ref, 3))
    with ra.HipRenderer(max_sample_count=3, max_bounces=12, seed=7, semantics=1) as r:
        img_vk = r.render_frame(sc)
        acc_vk = r.accumulator()
    ref_vk = oracle.vk_render(sc, ra.camera_vulkan_inverse_view(sc.camera), 3, 12, threads=8)
    np.testing.assert_array_equal(acc_vk.view(np.uint32), ref_vk.view(np.uint32))
    np.testing.assert_array_equal(img_vk.reshape(-1), oracle.vk_tonemap(ref_vk, 3))


@pytest.mark.parametrize("bvh4", [0, 1])
@pytest.mark.parametrize("lds_bvh", [0, 14336, 1 << 17])
def test_lds_resident_bvh_matches_oracle(lds_bvh, bvh4):
    # 97 objects (at most 9.3 KB of nodes + primitives): traversed from LDS
    # unless disabled, through the binary or the 4-wide tree
    sc = scenegen.plain_1080p(1, 96, 54)
    img, acc = _gpu_acc(sc, 3, 12, 21, lds_bvh=lds_bvh, bvh4=bvh4)
    ref = oracle.render(sc, 3, 12, 21)
    np.testing.assert_array_equal(acc.view(np.uint32), ref.view(np.uint32))
    np.testing.assert_array_equal(img.reshape(-1), oracle.tonemap(ref, 3))


@pytest.mark.parametrize("bvh4", [0, 1])
@pytest.mark.parametrize("leaf", [1, 4])
def test_clustered_scene_both_trees_match_oracle(leaf, bvh4):
    # 4096 clustered objects: a deep tree walked from global memory
    sc = scenegen.clustered(2, 4096, 160, 90)
    assert ra.host_bvh4_info(sc, leaf)[0] == abi.RD_OK
    img, acc = _gpu_acc(sc, 2, 12, 8, leaf=leaf, bvh4=bvh4)
    ref = oracle.render(sc, 2, 12, 8, threads=8)
    np.testing.assert_array_equal(acc.view(np.uint32), ref.view(np.uint32))


def test_default_rscn_matches_oracle():
    sc = Scene.load(os.path.join(GOLDEN, "default.rscn"))
    img, acc = _gpu_acc(sc, 4, 12, 11)
    ref = oracle.render(sc, 4, 12, 11)
    assert np.all(_close(acc, ref))
    assert np.mean(acc.view(np.uint32) == ref.view(np.uint32)) == 1.0


def test_progressive_samples_equal_full_frame():
    """render_sample one pass at a time (cpu.rs:142-158) == render_frame."""
    sc = scenegen.plain_small(0)
    with ra.HipRenderer(max_sample_count=6, max_bounces=8, seed=3) as r:
        full = r.render_frame(sc)
        acc_full = r.accumulator()
    with ra.HipRenderer(max_sample_count=6, max_bounces=8, seed=3) as r:
        r.new_frame(sc)
        r.set_samples_per_pass(4)
        frames = []
        while True:
            f = r.render_sample(sc)
            if f is None:
                break
            frames.append(f)
        assert len(frames) == 2 and r.sample_count == 6
        np.testing.assert_array_equal(r.accumulator(), acc_full)
        np.testing.assert_array_equal(frames[-1], full)
        prof = r.profiler()
        assert prof["frame_ns"] is not None and prof["render_ns"] is not None


class _Timer:
    """timing.rs:68-127 Timer (start / end / end_if_not_ended / end_multiple)."""

    def __init__(self):
        self.t0 = None
        self.duration = None

    def start(self):
        self.t0 = time.perf_counter()

    def start_if_not_started(self):
        if self.t0 is None:
            self.start()

    def end(self):
        if self.t0 is not None:
            self.duration = time.perf_counter() - self.t0

    def end_if_not_ended(self):
        if self.duration is None:
            self.end()

    def end_multiple(self, count):
        if self.t0 is not None:
            self.duration = (time.perf_counter() - self.t0) / count


def test_rust_shim_call_sequence_sets_every_timer():
    """raydar_amd/rust/renderer/hip.rs replayed call for call through ctypes:
    render_frame = new_frame (frame + prepare timers, rd_new_frame), then the
    timers of render_next_sample (cpu.rs:193-219) around one rd_render_sample
    pass of every sample, sample timer ended with end_multiple; render_sample =
    one sample per call.  print_profiling_metrics (main.rs:61-107) needs all
    four Rust-side durations; the library's own rd_profiler reports all four
    too, and the images equal the oracle's."""
    sc = scenegen.plain_small(0)
    spp = 5
    with ra.HipRenderer(max_sample_count=spp, max_bounces=12, seed=3) as r:
        prof = {k: _Timer() for k in ("frame", "prepare", "render", "sample")}
        # render_frame -> new_frame
        prof["frame"].start()
        prof["prepare"].start()
        r.new_frame(sc)
        remaining = r.max_sample_count - r.sample_count
        assert remaining == spp
        prof["prepare"].end_if_not_ended()
        prof["render"].start_if_not_started()
        prof["sample"].start()
        r.set_samples_per_pass(remaining)
        img = r.render_sample(sc)
        prof["render"].end()
        prof["frame"].end()
        prof["sample"].end_multiple(remaining)
        assert img is not None and r.sample_count == spp
        assert all(t.duration is not None for t in prof.values())
        lib_t = r.profiler()
        assert all(v is not None and v > 0 for v in lib_t.values()), lib_t
        acc = r.accumulator()
        r.set_samples_per_pass(1)
        assert r.render_sample(sc) is None  # cpu.rs:143-145
    ref = oracle.render(sc, spp, 12, 3)
    np.testing.assert_array_equal(acc.view(np.uint32), ref.view(np.uint32))
    np.testing.assert_array_equal(img.reshape(-1), oracle.tonemap(ref, spp))
    # the editor's loop: one render_sample per frame after new_frame
    with ra.HipRenderer(max_sample_count=3, max_bounces=12, seed=3) as r:
        prof = {k: _Timer() for k in ("frame", "prepare", "render", "sample")}
        prof["frame"].start()
        prof["prepare"].start()
        r.new_frame(sc)
        while r.sample_count < r.max_sample_count:
            prof["prepare"].end_if_not_ended()
            prof["render"].start_if_not_started()
            prof["sample"].start()
            assert r.render_sample(sc) is not None
            if r.sample_count == r.max_sample_count:
                prof["render"].end()
                prof["frame"].end()
            prof["sample"].end()
        assert all(t.duration is not None for t in prof.values())
        assert all(v is not None for v in r.profiler().values())
    # rd_render_frame alone: all four set, the sample timer is the per-sample average
    with ra.HipRenderer(max_sample_count=8, max_bounces=12, seed=3) as r:
        r.render_frame(sc)
        t = r.profiler()
        assert all(v is not None and v > 0 for v in t.values()), t
        assert t["sample_ns"] * 8 <= t["render_ns"] * 1.05 + 1e-4


def test_rmse_against_fp64_truth():
    """North-star image check at equal spp, averaged over 4 seeds."""
    z = np.load(os.path.join(GOLDEN, "plain_small_truth.npz"))
    sc = scenegen.plain_small(0)
    pix, truth, spp = z["pixels"], z["truth"], int(z["test_spp"])
    g_err, c_err = [], []
    for seed in range(4):
        _, acc = _gpu_acc(sc, spp, 12, 100 + seed)
        g = acc[pix, :3] / spp
        c = oracle.render(sc, spp, 12, 100 + seed, pixels=pix)[:, :3] / spp
        g_err.append(np.sqrt(np.mean((g - truth) ** 2)))
        c_err.append(np.sqrt(np.mean((c - truth) ** 2)))
    ratio = np.mean(g_err) / np.mean(c_err)
    print(f"RMSE gpu {np.mean(g_err):.6f} cpu {np.mean(c_err):.6f} ratio {ratio:.6f}")
    assert ratio <= 1.01


def test_empty_scene_is_sky_closed_form():
    cam = make_camera((0, 1, 5), (0, 0, 0), (0, 1, 0), 64, 48, fov_deg=60)
    sc = Scene(cam, sky_world((0.2, 0.4, 0.9), (1.0, 0.9, 0.8)), np.zeros(0, abi.OBJECT_DTYPE))
    _, acc = _gpu_acc(sc, 3, 12, 0)
    rays = oracle.camera_rays(sc, np.arange(64 * 48, dtype=np.uint32))
    sky = oracle.world(sc, rays[:, 3:])
    assert np.all(_close(acc[:, :3] / 3, sky, 1e-6))
    assert np.all(acc[:, 3] == 3)


def test_single_object_solid_world_closed_form():
    """One convex non-transmissive object under SolidColor(L): hit = albedo*L, miss = L."""
    cam = make_camera((0, 0, 6), (0, 0, 0), (0, 1, 0), 64, 64, fov_deg=40)
    L = (0.7, 0.5, 0.3)
    alb = (0.25, 0.5, 0.75)
    for kind in (abi.RD_SPHERE, abi.RD_CUBE):
        sc = Scene(cam, solid_world(L), np.array([make_object(kind, (0, 0, 0), 1.0, albedo=alb,
                                                               roughness=1.0)]))
        _, acc = _gpu_acc(sc, 4, 12, 1)
        rays = oracle.camera_rays(sc, np.arange(64 * 64, dtype=np.uint32))
        hit = oracle.trace(sc, rays)[0] >= 0
        v = acc[:, :3] / 4
        assert hit.sum() > 100
        assert np.allclose(v[hit], np.float32(alb) * np.float32(L), rtol=1e-6)
        assert np.allclose(v[~hit], np.float32(L), rtol=1e-6)


def test_enclosing_emitter_closed_form():
    """Camera inside an emissive sphere: every bounce re-hits it and adds
    emission unattenuated (cpu.rs:332-333), so pixel = max_bounces * E."""
    cam = make_camera((0, 0, 0.5), (0, 0, -1), (0, 1, 0), 32, 32, fov_deg=60)
    sc = Scene(cam, sky_world(), np.array([make_object(abi.RD_SPHERE, (0, 0, 0), 10.0,
                                                       albedo=(0.5, 0.5, 0.5), roughness=1.0,
                                                       emission_color=(1.0, 0.5, 0.25),
                                                       emission_strength=2.0)]))
    for bounces in (1, 5, 12):
        _, acc = _gpu_acc(sc, 2, bounces, 2)
        e = np.float32([2.0, 1.0, 0.5]) * bounces
        assert np.allclose(acc[:, :3] / 2, e, rtol=1e-6)


def test_zero_bounces_and_zero_samples():
    sc = scenegen.plain_small(0)
    img, acc = _gpu_acc(sc, 2, 0, 0)
    assert np.all(acc[:, :3] == 0) and np.all(acc[:, 3] == 2)
    assert np.all(img[..., :3] == 0) and np.all(img[..., 3] == 255)
    with ra.HipRenderer(max_sample_count=0, max_bounces=12) as r:
        img = r.render_frame(sc)
        assert np.all(img == 0)  # acc/0 = NaN -> 0 (cpu.rs:224)
        assert r.render_sample(sc) is None


def test_transparent_world_is_an_error():
    sc = scenegen.plain_small(0)
    sc.world = dict(sc.world, kind=abi.RD_WORLD_TRANSPARENT)
    with ra.HipRenderer(max_sample_count=1) as r:
        with pytest.raises(ra.RaydarError) as e:
            r.render_frame(sc)
        assert e.value.status == abi.RD_ERR_UNSUPPORTED


def _read_png(path):
    import struct
    import zlib
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat = 8, b""
    w = h = None
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        kind, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        if kind == b"IHDR":
            w, h = struct.unpack(">II", body[:8])
            assert body[8:10] == b"\x08\x06"
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = [raw[y * (4 * w + 1):(y + 1) * (4 * w + 1)] for y in range(h)]
    assert all(r[0] == 0 for r in rows)  # filter type None
    return np.frombuffer(b"".join(r[1:] for r in rows), np.uint8).reshape(h, w, 4)


def test_cli_renders_scene_file(tmp_path):
    """raydar_hip (main.rs counterpart): flags, .rscn input, PNG output."""
    import subprocess
    root = os.path.dirname(GOLDEN)
    cli = os.path.join(os.path.dirname(root), "raydar_amd", "raydar_hip")
    out = str(tmp_path / "out.png")
    p = subprocess.run([cli, "--max-sample-count", "3", "--max-bounces", "5", "-o", out,
                        os.path.join(GOLDEN, "default.rscn")], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stderr
    assert "Max Samples: 3" in p.stdout and "Total Frame Time:" in p.stdout
    img = _read_png(out)
    sc = Scene.load(os.path.join(GOLDEN, "default.rscn"))
    ref = oracle.tonemap(oracle.render(sc, 3, 5, 0), 3).reshape(sc.height, sc.width, 4)
    np.testing.assert_array_equal(img, ref)


def test_resolution_change_reallocates():
    a = scenegen.plain_small(0)
    b = scenegen.plain_small(0)
    b.camera = make_camera((0.0, 2.0, 8.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 96, 64, fov_deg=40.0)
    with ra.HipRenderer(max_sample_count=3, max_bounces=4) as r:
        r.new_frame(a)
        assert r.render_sample(a).shape == (256, 256, 4)
        out = r.render_sample(b)
        assert out.shape == (64, 96, 4)
        assert r.sample_count == 2


def test_copy_calibration_kernel():
    # the HBM-counter reference point used by the traffic profile
    ms = ra.debug_copy_calibration(1 << 26, 3)
    assert ms.shape == (3,) and (ms > 0).all()
    with pytest.raises(ra.RaydarError) as e:
        ra.debug_copy_calibration(1000, 1)  # not a multiple of 16
    assert e.value.status == abi.RD_ERR_INVALID_ARGUMENT


def test_wide_stack_scene_matches_oracle():
    # 70000 objects: the BVH has more than 65535 nodes, so the traversal
    # stack switches to 32-bit entries (plain renderer and batch alike)
    from raydar_amd.batch import BatchRenderer, Job
    sc = scenegen.clustered(31, 70000, 40, 24)
    assert ra.host_bvh_info(sc, leaf=1)[1] > 65535
    ref = oracle.render(sc, 1, 6, 3, threads=8)
    img, acc = _gpu_acc(sc, 1, 6, 3)
    np.testing.assert_array_equal(acc.view(np.uint32), ref.view(np.uint32))
    job = Job(sc, 1, 6, 3)
    with BatchRenderer(leaf=1) as br:
        br.render([job, Job(scenegen.plain_small(0), 1, 4, 0)])
    np.testing.assert_array_equal(job.acc.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("opts,reuse_pipes", [(dict(pipes=1, queue=16384), 1),
                                              (dict(pipes=2, queue=16384), 2),
                                              (dict(lds_bvh=0, bvh4=0, pipes=1, queue=16384), 1),
                                              (dict(queue=4096), None), (dict(), None),
                                              (dict(pipes=1, queue=16384, hit_reuse=0), None)])
def test_camera_ray_hits_reused_after_each_pipes_first_sample(opts, reuse_pipes):
    # a pixel's camera ray does not depend on the sample (cpu.rs:234-251): a
    # pipe whose pass needs several fills of its queue traces one sample's
    # camera rays and gives its later camera paths their pixel's recorded
    # hit; a queue smaller than one sample, or a pass that fits in one fill
    # (41472 paths in a 16 Mi queue), traces them all
    sc = scenegen.plain_1080p(1, 96, 54)
    npix, spp = sc.width * sc.height, 8
    with ra.HipRenderer(max_sample_count=spp, max_bounces=12, seed=4, **opts) as r:
        r.reset_counters()
        r.set_option("count", 1)
        r.render_frame(sc)
        c = r.counters()
        acc = r.accumulator()
    assert c["camera_rays"] == spp * npix
    want = 0 if reuse_pipes is None else (spp - reuse_pipes) * npix
    assert c["reused_hits"] == want
    assert c["sphere_hits"] + c["cube_hits"] + c["misses"] == c["segments"] + c["reused_hits"]
    ref = oracle.render(sc, spp, 12, 4)
    np.testing.assert_array_equal(acc.view(np.uint32), ref.view(np.uint32))


def test_camera_ray_hits_carry_across_a_frames_passes():
    # progressive rendering, one sample per call: the frame's first pass
    # records every pixel's camera-ray hit and its later passes reuse them;
    # restart_frame and new_frame start a frame that records them again
    sc = scenegen.plain_1080p(1, 96, 54)
    npix, spp = sc.width * sc.height, 5
    with ra.HipRenderer(max_sample_count=spp, max_bounces=12, seed=4) as r:
        r.reset_counters()
        r.set_option("count", 1)
        r.new_frame(sc)
        for _ in range(spp):
            r.render_sample(sc)
        c = r.counters()
        acc = r.accumulator()
        assert c["reused_hits"] == (spp - 1) * npix
        ref = oracle.render(sc, spp, 12, 4)
        np.testing.assert_array_equal(acc.view(np.uint32), ref.view(np.uint32))
        for start in (r.restart_frame, lambda: r.new_frame(scenegen.plain_1080p(2, 96, 54))):
            start()
            r.reset_counters()
            r.render_sample()
            assert r.counters()["reused_hits"] == 0
            r.render_sample()
            assert r.counters()["reused_hits"] == npix
        acc = r.accumulator()
    ref = oracle.render(scenegen.plain_1080p(2, 96, 54), 2, 12, 4)
    np.testing.assert_array_equal(acc.view(np.uint32), ref.view(np.uint32))


def test_event_counts_agree_between_trees():
    # the counting pass behind the roofline figure: the two trees give the
    # same paths, so every per-path event agrees and only the traversal
    # work (boxes and primitives tested) differs
    sc = scenegen.clustered(2, 4096, 160, 90)
    counts = {}
    for bvh4 in (0, 1):
        with ra.HipRenderer(max_sample_count=2, max_bounces=12, seed=8, bvh4=bvh4) as r:
            r.reset_counters()
            r.set_option("count", 1)
            r.render_frame(sc)
            counts[bvh4] = r.counters()
    a, b = counts[0], counts[1]
    assert a["camera_rays"] == b["camera_rays"] == sc.width * sc.height * 2
    for k in ("sphere_hits", "cube_hits", "glass_scatters", "misses", "accumulates"):
        assert a[k] == b[k], k
    # every shaded path got its hit from a query or from its pixel's camera
    # ray (the second sample of this 2-spp frame reuses the first's)
    assert a["segments"] == b["segments"] and a["reused_hits"] == b["reused_hits"]
    assert a["sphere_hits"] + a["cube_hits"] + a["misses"] == a["segments"] + a["reused_hits"]
    for k in ("box_tests", "sphere_tests", "cube_tests"):
        assert a[k] > 0 and b[k] > 0, k
    # the lanes kernel's SIMD-efficiency counts obey their definitions
    for c in (a, b):
        assert 0 < c["wave_steps"] <= c["lane_steps"] <= 64 * c["wave_steps"]
        assert 0 < c["prim_trips"] and c["cube_trips"] <= c["prim_trips"]
        assert c["prim_trips"] <= c["sphere_tests"] + c["cube_tests"] <= 64 * c["prim_trips"]
        assert c["sort_steps"] <= c["wave_steps"] and c["refills"] > 0
    # a binary step tests two boxes per lane, a 4-wide step up to four
    assert a["box_tests"] == 2 * a["lane_steps"]
    assert b["box_tests"] <= 4 * b["lane_steps"] and b["sort_steps"] > 0 == a["sort_steps"]


@pytest.mark.parametrize("spp,pipes,queue,lds,used", [
    (3, 2, 4096, 14336, 2), (5, 4, 4096, 14336, 4), (2, 4, 4096, 14336, 2), (1, 2, 4096, 14336, 1),
    (5, 4, 1 << 16, 14336, 2), (5, 2, 1 << 23, 14336, 1),
    (5, 0, 4096, 14336, 2),  # automatic: the tree is traversed from LDS
    (5, 0, 4096, 0, 3)])     # automatic: the tree stays in global memory
def test_pass_divided_over_pipes_matches_oracle(spp, pipes, queue, lds, used):
    # a pass of spp samples runs on up to min(pipes, spp) streams, each
    # tracing a contiguous run of samples into its own radiance slots and
    # each given at least half a queue of paths (160x90 = 14400 per sample)
    sc = scenegen.plain_1080p(1, 160, 90)
    with ra.HipRenderer(max_sample_count=spp, max_bounces=12, seed=4, pipes=pipes,
                        queue=queue, lds_bvh=lds) as r:
        img = r.render_frame(sc)
        acc = r.accumulator()
        assert r.get_option("pipes_used") == used
    ref = oracle.render(sc, spp, 12, 4)
    np.testing.assert_array_equal(acc.view(np.uint32), ref.view(np.uint32))
    np.testing.assert_array_equal(img.reshape(-1), oracle.tonemap(ref, spp))
    with ra.HipRenderer(max_sample_count=spp, max_bounces=12, seed=4) as r:
        assert r.get_option("pipes") == 0  # automatic
        for bad in (-1, 5):
            with pytest.raises(ra.RaydarError) as e:
                r.set_option("pipes", bad)
            assert e.value.status == abi.RD_ERR_INVALID_ARGUMENT


def test_passes_over_pipes_share_the_radiance_arena_in_order():
    # a 65536-slot arena holds 4 samples of 160x90: 12 spp run as three
    # passes, each divided over four pipes; the pipes of a pass write the
    # arena slots the previous pass's accumulation reads, so they must start
    # after it (an event on the main stream)
    sc = scenegen.plain_1080p(1, 160, 90)
    with ra.HipRenderer(max_sample_count=12, max_bounces=12, seed=6, pipes=4, queue=4096,
                        l_capacity=1 << 16) as r:
        img = r.render_frame(sc)
        acc = r.accumulator()
        assert r.get_option("pipes_used") == 4
    ref = oracle.render(sc, 12, 12, 6)
    np.testing.assert_array_equal(acc.view(np.uint32), ref.view(np.uint32))
    np.testing.assert_array_equal(img.reshape(-1), oracle.tonemap(ref, 12))


def test_seed_and_bounce_changes_between_frames_on_one_renderer():
    # the job record on the device is re-uploaded only when the seed or the
    # bounce limit changed: every frame must match the oracle for the
    # settings it was rendered with, also after a scene of another size
    small = scenegen.plain_small(0)
    other = scenegen.plain_1080p(1, 96, 54)
    with ra.HipRenderer(max_sample_count=2, max_bounces=12, seed=1) as r:
        steps = [(small, 1, 12, None), (small, 2, 12, "seed"), (small, 2, 4, "bounces"),
                 (small, 2, 4, None), (other, 2, 4, None), (other, 7, 9, "both")]
        for sc, seed, bounces, change in steps:
            if change in ("seed", "both"):
                r.set_seed(seed)
            if change in ("bounces", "both"):
                r.max_bounces = bounces
            img = r.render_frame(sc)
            ref = oracle.render(sc, 2, bounces, seed)
            np.testing.assert_array_equal(r.accumulator().view(np.uint32), ref.view(np.uint32))
            np.testing.assert_array_equal(img.reshape(-1), oracle.tonemap(ref, 2))


def _host_threads():
    return max(1, min(16, os.cpu_count() or 1))


def test_full_size_plain_1080p_matches_oracle():
    # the 1080p configuration at full resolution: 5 spp is 10.4 M paths,
    # enough for two pipes at the default 8 M-path queue
    sc = scenegen.plain_1080p(1)
    with ra.HipRenderer(max_sample_count=5, max_bounces=12, seed=7) as r:
        img = r.render_frame(sc)
        acc = r.accumulator()
        assert r.get_option("pipes_used") == 2
    ref = oracle.render(sc, 5, 12, 7, threads=_host_threads())
    np.testing.assert_array_equal(acc.view(np.uint32), ref.view(np.uint32))
    np.testing.assert_array_equal(img.reshape(-1), oracle.tonemap(ref, 5))


def test_full_size_clustered_1080p_matches_oracle_on_pixels():
    # 4096 clustered primitives at full resolution, 8 spp (three pipes, tree
    # in global memory); the oracle's linear scan checks every 257th pixel
    sc = scenegen.clustered(2)
    with ra.HipRenderer(max_sample_count=8, max_bounces=12, seed=9) as r:
        r.render_frame(sc)
        acc = r.accumulator()
        assert r.get_option("pipes_used") == 3
    pix = np.arange(0, sc.width * sc.height, 257, dtype=np.uint32)
    ref = oracle.render(sc, 8, 12, 9, pixels=pix, threads=_host_threads())
    np.testing.assert_array_equal(acc[pix].view(np.uint32), ref.view(np.uint32))
