"""GPU parity of each hot-path stage against the oracle, on identical inputs.

The three checks of the path's contract:
  1. closest hit (src/renderer/cpu.rs:344-394): object id bit-exact for
     identical rays; hit distance and normal within 1e-5 relative.
  2. scatter (cpu.rs:262-333): directions and attenuations within 1e-5 for
     identical draw words.
  3. whole images: see test_gpu_render.py.
Each test also reports how many values were bit-identical (the kernels run
the oracle's FP32 operation sequence, so all of them normally are).
"""
import zlib

import numpy as np
import pytest

import oracle
import raydar_amd as ra
from raydar_amd import scenegen
from raydar_amd.scene import Scene
from tests import raygen

pytestmark = pytest.mark.gpu

REL = 1e-5


def _close(a, b, rel=REL):
    # equal values (hits at infinity included) and NaN on both sides (the
    # normal at such a hit) match; everything else within rel
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = np.maximum(np.maximum(np.abs(a), np.abs(b)), 1.0)
    with np.errstate(invalid="ignore"):
        near = np.abs(a - b) <= rel * scale
    return near | (a == b) | (np.isnan(a) & np.isnan(b))


material_zoo = raygen.material_zoo

SCENES = {
    "plain_small": lambda: scenegen.plain_small(0),
    "plain_1080p": lambda: scenegen.plain_1080p(1),
    "clustered_512": lambda: scenegen.clustered(2, 512, 640, 360),
    "default_rscn": lambda: Scene.load(__import__("os").path.join(
        __import__("os").path.dirname(__file__), "golden", "default.rscn")),
    "benchmark_rscn": lambda: Scene.load(__import__("os").path.join(
        __import__("os").path.dirname(__file__), "golden", "benchmark.rscn")),
    "material_zoo": material_zoo,
}


def test_philox_device_matches_oracle_and_kat():
    rng = np.random.default_rng(0)
    inp = rng.integers(0, 2**32, (4096, 6), dtype=np.uint64).astype(np.uint32)
    kat = np.array([[0, 0, 0, 0, 0, 0],
                    [0xffffffff] * 6,
                    [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344, 0xa4093822, 0x299f31d0]],
                   np.uint32)
    inp = np.concatenate([kat, inp])
    out = ra.debug_philox(inp)
    ref = np.array([oracle.philox(r[:4], r[4:]) for r in inp])
    np.testing.assert_array_equal(out, ref)
    # Random123 known-answer vectors for philox4x32-10
    np.testing.assert_array_equal(out[0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8])
    np.testing.assert_array_equal(out[1], [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd])
    np.testing.assert_array_equal(out[2], [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1])


@pytest.mark.parametrize("name", sorted(SCENES))
def test_camera_rays_bit_exact(name):
    sc = SCENES[name]()
    rng = np.random.default_rng(1)
    pix = rng.integers(0, sc.width * sc.height, 20000).astype(np.uint32)
    pix[:4] = [0, sc.width - 1, sc.width * (sc.height - 1), sc.width * sc.height - 1]
    g = ra.debug_camera_rays(sc, pix)
    o = oracle.camera_rays(sc, pix)
    assert np.all(_close(g, o))
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32)), \
        f"{(g.view(np.uint32) != o.view(np.uint32)).sum()} camera-ray words differ in the last bits"


@pytest.mark.parametrize("name", sorted(SCENES))
@pytest.mark.parametrize("leaf,brute", [(1, False), (4, False), (8, False), (4, True)])
def test_closest_hit_parity(name, leaf, brute):
    sc = SCENES[name]()
    rays = raygen.mixed(sc, 3000, seed=zlib.crc32(f"{name}/{leaf}".encode()))
    ids_g, ts_g, n_g, f_g = ra.debug_trace(sc, rays, leaf=leaf, brute=brute)
    ids_o, ts_o, n_o, f_o = oracle.trace(sc, rays)
    np.testing.assert_array_equal(ids_g, ids_o)
    hit = ids_o >= 0
    assert np.all(_close(ts_g[hit], ts_o[hit]))
    assert np.all(_close(n_g[hit], n_o[hit]))
    np.testing.assert_array_equal(f_g[hit], f_o[hit])
    exact = np.mean(ts_g[hit].view(np.uint32) == ts_o[hit].view(np.uint32)) if hit.any() else 1.0
    print(f"{name} leaf={leaf} brute={brute}: {len(rays)} rays, hit rate {hit.mean():.3f}, "
          f"t bit-exact {exact:.6f}")
    assert exact == 1.0


# Forms of the closest-hit kernel the renderer launches (Engine::run_tickets for
# one scene, the streamers for batches and the server): tree width, LDS or
# global tree, 16- or 32-bit stack, single- or multi-job, persistent lanes or
# one ray per lane, and a small grid so that every wave refills its lanes.
EXTEND_VARIANTS = {
    "w4_lds": dict(bvh4=1, lds=1),
    "w4_global": dict(bvh4=1, lds=0),
    "bin_lds": dict(bvh4=0, lds=1),
    "bin_global16": dict(bvh4=0, wide=0),
    "bin_global32": dict(bvh4=0, wide=1),
    "multi_w4": dict(bvh4=1, jobs=3),
    "multi_bin32": dict(bvh4=0, wide=1, jobs=3),
    "multi_bin16_noref": dict(bvh4=0, jobs=3, refill=0),
    "single_bin32_noref": dict(bvh4=0, wide=1, refill=0),
    "w4_refill1_grid3": dict(bvh4=1, refill=1, block=64, grid=3),
    "multi_w4_refill16_grid5": dict(bvh4=1, jobs=5, refill=16, block=128, grid=5),
    "w4_lds_leaf4_grid2": dict(bvh4=1, lds=1, leaf=4, block=64, grid=2),
    # leaf 1 (the default above) takes the one-primitive-per-leaf kernels; leaf 2 the general ones
    "w4_global_leaf2": dict(bvh4=1, lds=0, leaf=2),
    # the queue dealt to the waves in chunks (render option lane_chunk)
    "w4_lds_chunk64": dict(bvh4=1, lds=1, lane_chunk=64),
    "w4_global_chunk64_grid7": dict(bvh4=1, lds=0, lane_chunk=64, grid=7),
    "multi_w4_chunk128": dict(bvh4=1, jobs=3, lane_chunk=128),
    "bin_global32_chunk256": dict(bvh4=0, wide=1, lane_chunk=256),
}


@pytest.mark.parametrize("variant", sorted(EXTEND_VARIANTS))
@pytest.mark.parametrize("name", sorted(SCENES))
def test_production_extend_parity(name, variant):
    """The adversarial ray sets through the kernel that renders (k_extend_lanes
    / k_extend), not the debug walk: ids bit-exact, t and normals within 1e-5
    (and in fact bit-exact), against the oracle's linear scan (cpu.rs:344-394)."""
    sc = SCENES[name]()
    opts = dict(EXTEND_VARIANTS[variant])
    leaf = opts.pop("leaf", 1)
    rays = raygen.mixed(sc, 1500, seed=zlib.crc32(f"extend/{name}/{variant}".encode()))
    got = ra.debug_extend(sc, rays, leaf=leaf, **opts)
    if got is None:
        pytest.skip(f"the renderer does not launch the {variant} form for {name}")
    ids_g, ts_g, n_g, f_g = got
    ids_o, ts_o, n_o, f_o = oracle.trace(sc, rays)
    np.testing.assert_array_equal(ids_g, ids_o)
    hit = ids_o >= 0
    assert np.all(_close(ts_g[hit], ts_o[hit]))
    assert np.all(_close(n_g[hit], n_o[hit]))
    np.testing.assert_array_equal(f_g[hit], f_o[hit])
    exact = np.mean(ts_g[hit].view(np.uint32) == ts_o[hit].view(np.uint32)) if hit.any() else 1.0
    print(f"{name} {variant}: {len(rays)} rays, hit rate {hit.mean():.3f}, t bit-exact {exact:.6f}")
    assert exact == 1.0


def test_production_extend_forms_cover_the_renderer():
    """Each kernel form of the parametrised test runs on at least one scene
    (a form every scene skips would leave that kernel untested)."""
    sc = scenegen.plain_small(0)
    rays = raygen.mixed(sc, 50, seed=1)
    for variant, opts in EXTEND_VARIANTS.items():
        opts = dict(opts)
        leaf = opts.pop("leaf", 1)
        assert ra.debug_extend(sc, rays, leaf=leaf, **opts) is not None, variant


_scatter_inputs = raygen.scatter_inputs


@pytest.mark.parametrize("name", ["plain_small", "plain_1080p", "default_rscn", "material_zoo",
                                  "benchmark_rscn"])
def test_scatter_parity(name):
    sc = SCENES[name]()
    packed, bits = _scatter_inputs(sc, 6000, 7)
    g = ra.debug_scatter(sc, packed, bits)
    o = oracle.scatter(sc, packed, bits)
    np.testing.assert_array_equal(g[:, 12], o[:, 12])  # transmission branch taken
    assert np.all(_close(g, o)), f"max |diff| {np.max(np.abs(g - o))}"
    exact = np.mean(g.view(np.uint32) == o.view(np.uint32))
    print(f"{name}: {len(packed)} scatters, {o[:, 12].mean():.3f} transmissive, bit-exact {exact:.6f}")
    assert exact == 1.0


def test_scatter_roughness_and_degenerate_pins_on_device():
    """The closed-form cases of tests/pins.py on the GPU's scatter."""
    from tests import pins
    assert pins.failures(ra.debug_scatter) == []


@pytest.mark.parametrize("name", ["plain_small", "default_rscn"])
def test_world_sample_parity(name):
    sc = SCENES[name]()
    rng = np.random.default_rng(5)
    d = rng.normal(size=(5000, 3)).astype(np.float32) * rng.uniform(0.1, 3, (5000, 1)).astype(np.float32)
    d[:3] = [[0, 1, 0], [0, -1, 0], [1, 0, 0]]
    g = ra.debug_world(sc, d)
    o = oracle.world(sc, d)
    np.testing.assert_array_equal(g.view(np.uint32), o.view(np.uint32))
