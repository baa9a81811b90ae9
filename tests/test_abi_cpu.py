"""CPU checks of the C-ABI library: it loads, exports every declared symbol,
fails cleanly without a GPU, and its host-side parts (the .rscn reader and
writer, the BVH builder and its traversal rule) agree with the oracle."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import oracle
import raydar_amd as ra
from raydar_amd import abi, scenegen
from raydar_amd.scene import Scene, make_camera, make_object, sky_world
from tests import raygen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(rd_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = C.CDLL(ra.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_config_default_matches_renderer_config():
    cfg = abi.RdConfig()
    ra.check(ra.lib().rd_config_default(C.byref(cfg)))
    assert (cfg.max_sample_count, cfg.max_bounces) == (1024, 12)  # renderer/mod.rs:16-23


def test_create_without_gpu_fails_cleanly():
    import subprocess
    import sys
    code = ("import raydar_amd as ra\n"
            "try:\n    ra.HipRenderer()\n    print('CREATED')\n"
            "except ra.RaydarError as e:\n    print('ERR', e.status, e)\n")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True,
                         text=True, timeout=120).stdout
    assert out.startswith("ERR 4"), out
    assert "HIP device" in out or "hip" in out.lower()


@pytest.mark.parametrize("name", ["default.rscn", "benchmark.rscn"])
def test_native_rscn_reader_matches_python(name):
    path = os.path.join(GOLDEN, name)
    a, b = Scene.load(path), ra.load_rscn_native(path)
    for k in a.camera:
        assert np.array_equal(np.float32(a.camera[k]), np.float32(b.camera[k])), k
    assert a.objects.tobytes() == b.objects.tobytes()
    assert {k: np.float32(v).tolist() for k, v in a.world.items()} == \
        {k: np.float32(v).tolist() for k, v in b.world.items()}


def test_benchmark_rscn_contents():
    sc = Scene.load(os.path.join(GOLDEN, "benchmark.rscn"))
    kinds = sc.objects["kind"]
    assert (sc.width, sc.height) == (3840, 2160)
    assert (kinds == abi.RD_SPHERE).sum() == 70 and (kinds == abi.RD_CUBE).sum() == 113


@pytest.mark.parametrize("name", ["default.rscn", "benchmark.rscn"])
def test_rscn_round_trip(tmp_path, name):
    src = Scene.load(os.path.join(GOLDEN, name))
    p1 = str(tmp_path / "native.rscn")
    ra.save_rscn_native(src, p1)
    back = ra.load_rscn_native(p1)
    assert back.objects.tobytes() == src.objects.tobytes()
    for k in src.camera:
        assert np.array_equal(np.float32(src.camera[k]), np.float32(back.camera[k])), k
    p2 = str(tmp_path / "python.rscn")
    src.save(p2)
    assert Scene.load(p2).objects.tobytes() == src.objects.tobytes()
    assert ra.load_rscn_native(p2).objects.tobytes() == src.objects.tobytes()


@pytest.mark.parametrize("text,status", [
    ("{", abi.RD_ERR_PARSE),
    ('{"camera": {}}', abi.RD_ERR_PARSE),
    ("[1, 2]", abi.RD_ERR_PARSE),
    ('{"camera": 1, "world": "Transparent", "objects": []} x', abi.RD_ERR_PARSE),
])
def test_rscn_parse_errors(text, status):
    L = ra.lib()
    p = C.POINTER(abi.RdScene)()
    st = L.rd_scene_parse_rscn(text.encode(), len(text), C.byref(p))
    assert st == status
    assert ra.last_error()


def test_rscn_missing_file():
    p = C.POINTER(abi.RdScene)()
    assert ra.lib().rd_scene_load_rscn(b"/nonexistent/x.rscn", C.byref(p)) == abi.RD_ERR_IO


def test_transparent_world_parses(tmp_path):
    sc = scenegen.plain_small(0)
    sc.world = dict(sc.world, kind=abi.RD_WORLD_TRANSPARENT)
    p = str(tmp_path / "t.rscn")
    sc.save(p)
    assert ra.load_rscn_native(p).world["kind"] == abi.RD_WORLD_TRANSPARENT


# ------------------------------------------------------------------ BVH ----
def _random_scene(seed, n):
    rng = np.random.default_rng(seed)
    objs = [make_object(int(rng.integers(0, 2)), rng.uniform(-20, 20, 3), float(rng.uniform(0.01, 3)))
            for _ in range(n)]
    return Scene(make_camera((0, 5, 40), (0, 0, 0), (0, 1, 0), 64, 64, fov_deg=50), sky_world(),
                 np.array(objs))


@pytest.mark.parametrize("leaf", [1, 2, 4, 8])
@pytest.mark.parametrize("n", [1, 2, 7, 100, 2048])
def test_bvh_invariants(leaf, n):
    sc = _random_scene(n, n)
    st, nodes, depth, err = ra.host_bvh_info(sc, leaf)
    assert st == abi.RD_OK, err
    assert nodes >= 1 and depth <= 38
    st, nodes4, stack4, err = ra.host_bvh4_info(sc, leaf)
    assert st == abi.RD_OK, err
    # every 4-wide node absorbs at least one binary node
    assert 1 <= nodes4 <= nodes and stack4 <= 3 * depth


def test_bvh_degenerate_inputs():
    same = Scene(make_camera((0, 0, 9), (0, 0, 0), (0, 1, 0), 8, 8, fov_deg=40), sky_world(),
                 np.array([make_object(abi.RD_SPHERE, (1, 2, 3), 0.5)] * 300))
    st, _, depth, err = ra.host_bvh_info(same, 1)
    assert st == abi.RD_OK, err
    ids, _ = ra.host_trace(same, np.array([[1, 2, 9, 0, 0, -1]], np.float32), 1)
    assert ids[0] == 0  # 300 coincident spheres: the lowest index wins
    ids, _ = ra.host_trace(same, np.array([[1, 2, 9, 0, 0, -1]], np.float32), 1, 4)
    assert ids[0] == 0
    empty = Scene(same.camera, sky_world(), np.zeros(0, abi.OBJECT_DTYPE))
    ids, _ = ra.host_trace(empty, np.array([[0, 0, 9, 0, 0, -1]], np.float32))
    assert ids[0] == -1


@pytest.mark.parametrize("width", [2, 4])
@pytest.mark.parametrize("scene_name", ["plain_small", "plain_1080p", "clustered_1024", "benchmark"])
@pytest.mark.parametrize("leaf", [1, 4, 8])
def test_bvh_traversal_matches_linear_scan(scene_name, leaf, width):
    """The BVH culling rule (padded boxes, cull_limit) never drops the scan's
    answer, phantom hits of the FP32 sphere test included; the 4-wide walk
    (sorted children, fixed stack writes) also stays within its stack bound."""
    sc = {"plain_small": lambda: scenegen.plain_small(0),
          "plain_1080p": lambda: scenegen.plain_1080p(1),
          "clustered_1024": lambda: scenegen.clustered(2, 1024, 320, 180),
          "benchmark": lambda: Scene.load(os.path.join(GOLDEN, "benchmark.rscn"))}[scene_name]()
    rays = raygen.mixed(sc, 2500, seed=leaf)
    ids_o, ts_o, _, _ = oracle.trace(sc, rays)
    ids_h, ts_h = ra.host_trace(sc, rays, leaf, width)
    np.testing.assert_array_equal(ids_h, ids_o)
    hit = ids_o >= 0
    np.testing.assert_array_equal(ts_h[hit].view(np.uint32), ts_o[hit].view(np.uint32))


@pytest.mark.parametrize("scene_name", ["plain_small", "plain_1080p", "default", "zoo", "benchmark"])
def test_kernel_scatter_host_build_matches_oracle(scene_name):
    """pathmath.hpp's scatter (branch chosen first, only the chosen direction
    computed, draws generated lazily) equals the oracle's eager restatement
    of cpu.rs:262-333 bit for bit."""
    sc = {"plain_small": lambda: scenegen.plain_small(0),
          "plain_1080p": lambda: scenegen.plain_1080p(1),
          "default": lambda: Scene.load(os.path.join(GOLDEN, "default.rscn")),
          "zoo": raygen.material_zoo,
          "benchmark": lambda: Scene.load(os.path.join(GOLDEN, "benchmark.rscn"))}[scene_name]()
    packed, bits = raygen.scatter_inputs(sc, 4000, 21)
    h = ra.host_scatter(sc, packed, bits)
    o = oracle.scatter(sc, packed, bits)
    np.testing.assert_array_equal(h.view(np.uint32), o.view(np.uint32))


def test_kernel_scatter_host_build_meets_roughness_pins():
    from tests import pins
    assert pins.failures(ra.host_scatter) == []


@pytest.mark.parametrize("scene_name", ["plain_small", "default", "benchmark"])
def test_kernel_camera_host_build_matches_oracle(scene_name):
    sc = {"plain_small": lambda: scenegen.plain_small(0),
          "default": lambda: Scene.load(os.path.join(GOLDEN, "default.rscn")),
          "benchmark": lambda: Scene.load(os.path.join(GOLDEN, "benchmark.rscn"))}[scene_name]()
    pix = np.random.default_rng(2).integers(0, sc.width * sc.height, 5000).astype(np.uint32)
    np.testing.assert_array_equal(ra.host_camera_rays(sc, pix).view(np.uint32),
                                  oracle.camera_rays(sc, pix).view(np.uint32))


def test_nonfinite_rays_use_the_scan():
    sc = scenegen.plain_1080p(1)
    rays = raygen.mixed(sc, 50, seed=9)
    rays[::3, 3] = np.nan
    rays[1::3, 4] = np.inf
    ids_o, ts_o, _, _ = oracle.trace(sc, rays)
    ids_h, ts_h = ra.host_trace(sc, rays, 4)
    np.testing.assert_array_equal(ids_h, ids_o)


def test_radiance_ring_wrap_to_tail_is_full():
    # [0,40) and [40,80) of 100; release the first; [0,40) again ends exactly
    # at the oldest live range, so the ring is full: nothing more fits
    out = ra.ring_ops(100, [40, 40, -1, 40, 10, 20])
    assert list(out[:2]) == [0, 40] and out[3] == 0
    assert out[4] == -1 and out[5] == -1
    # releasing the range at 40 frees everything behind the newest one
    out = ra.ring_ops(100, [40, 40, -1, 40, -2, 60, 1])
    assert out[5] == 40 and out[6] == -1


def test_radiance_ring_never_overlaps_live_ranges():
    rng = np.random.default_rng(5)
    for cap in (7, 64, 1000):
        ops, sizes, live = [], [], set()
        for _ in range(3000):
            if live and rng.random() < 0.45:
                t = int(rng.choice(sorted(live)))
                live.discard(t)
                ops.append(-t - 1)
            else:
                n = int(rng.integers(1, cap // 2 + 2))
                ops.append(n)
                sizes.append(n)
                live.add(len(sizes) - 1)  # refused ones are dropped below
        out = ra.ring_ops(cap, ops)
        # replay: live ranges (reserved, not released) must never overlap,
        # and a reservation is refused only if the ring is not empty
        ranges, t, alive = {}, 0, set()
        for op, o in zip(ops, out):
            if op >= 0:
                if o >= 0:
                    assert 0 <= o and o + op <= cap
                    for a, b in (ranges[x] for x in alive):
                        assert o + op <= a or b <= o, (cap, o, op, a, b)
                    ranges[t] = (o, o + op)
                    alive.add(t)
                else:
                    assert alive or op > cap
                t += 1
            else:
                alive.discard(-op - 1)


def test_reused_buffer_size_classes():
    # a class never rounds a request down, wastes at most an eighth of it
    # above 4 KB, is its own class (reuse finds it again) and grows with the
    # request; powers of two are exact
    rng = np.random.default_rng(3)
    req = np.concatenate([np.arange(1, 20000, 7), rng.integers(1, 1 << 40, 20000),
                          (1 << np.arange(12, 41)), (1 << np.arange(12, 41)) + 1]).astype(np.uint64)
    req.sort()
    cls = ra.cache_classes(req)
    assert np.all(cls >= req)
    assert np.all(cls[req <= 4096] == 4096)
    big = req > 4096
    assert np.all(cls[big].astype(np.float64) <= req[big].astype(np.float64) * 1.125)
    assert np.array_equal(ra.cache_classes(cls), cls)
    assert np.all(np.diff(cls.astype(np.int64)) >= 0)
    pw = (1 << np.arange(12, 41)).astype(np.uint64)
    assert np.array_equal(ra.cache_classes(pw), pw)
    # eight classes per power of two
    assert len(np.unique(ra.cache_classes(np.arange(65537, 131073, dtype=np.uint64)))) == 8


def test_output_image_buffer_is_reused_or_rejected():
    # render_frame / render_sample write into `out` when given (a caller's
    # frame buffer reused across frames) and refuse buffers of another size,
    # type or layout before any device work
    sc = scenegen.plain_small(0)
    buf = np.zeros((sc.height, sc.width, 4), np.uint8)
    assert ra.HipRenderer._image(sc, buf) is buf
    assert ra.HipRenderer._image(sc, None).shape == (sc.height, sc.width, 4)
    for bad in (np.zeros((sc.height, sc.width, 3), np.uint8), np.zeros((sc.height, sc.width, 4), np.float32),
                np.zeros((sc.width, sc.height * 8), np.uint8)[:, ::2]):
        with pytest.raises(ValueError):
            ra.HipRenderer._image(sc, bad)


# ---------------------------------------------------------------- camera ---
_MATRICES = ("view_matrix", "proj_matrix", "inverse_view_matrix", "inverse_proj_matrix")


@pytest.mark.parametrize("name", ["default.rscn", "benchmark.rscn"])
def test_camera_update_matrices_reproduce_the_reference_scene_files(name):
    """rd_camera_update_matrices (camera.rs:210-231 in cgmath's FP32 order)
    recomputes the four matrices cgmath stored in the reference's scene files
    bit for bit: look_at_lh, perspective and both inverses."""
    sc = Scene.load(os.path.join(GOLDEN, name))
    cam = ra.camera_update_matrices(sc.camera)
    for k in _MATRICES:
        got = np.array(cam[k], np.float32)
        ref = np.array(sc.camera[k], np.float32)
        np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32), err_msg=k)


def test_camera_update_matrices_closed_forms_and_errors():
    base = make_camera((1.0, 2.0, 3.0), (0.5, 0.0, -1.0), (0, 1, 0), 640, 360, fov_deg=50)
    # orthographic: cgmath ortho(-s*a, s*a, -s, s, n, f) (camera.rs:218-225), whose
    # inverse composes to the identity
    ortho = dict(base, projection=abi.RD_PROJECTION_ORTHOGRAPHIC, projection_param=4.0)
    cam = ra.camera_update_matrices(ortho)
    a = np.float32(640) / np.float32(360)
    p = np.array(cam["proj_matrix"], np.float64).reshape(4, 4)  # p[c][r]
    assert np.isclose(p[0, 0], 2 / (2 * 4 * a), rtol=1e-6) and np.isclose(p[1, 1], 2 / 8, rtol=1e-6)
    assert np.isclose(p[2, 2], -2 / (1000 - 0.01), rtol=1e-6)
    assert np.isclose(p[3, 2], -(1000 + 0.01) / (1000 - 0.01), rtol=1e-6) and p[3, 3] == 1
    for m, mi in (("proj_matrix", "inverse_proj_matrix"), ("view_matrix", "inverse_view_matrix")):
        M = np.array(cam[m], np.float64).reshape(4, 4).T
        Mi = np.array(cam[mi], np.float64).reshape(4, 4).T
        assert np.allclose(M @ Mi, np.eye(4), atol=2e-6)
    # perspective and view against the double-precision formulas
    cam = ra.camera_update_matrices(base)
    ref = make_camera((1.0, 2.0, 3.0), (0.5, 0.0, -1.0), (0, 1, 0), 640, 360, fov_deg=50)
    for k in _MATRICES:
        assert np.allclose(cam[k], ref[k], rtol=2e-6, atol=2e-6), k
    # parameters the reference panics on: an error, the camera unchanged
    for bad in (dict(resolution_y=0), dict(projection_param=0.0), dict(projection_param=180.0),
                dict(near_clip=0.0), dict(far_clip=0.005), dict(projection=7),
                dict(target=[1.0, 2.0, 3.0])):  # eye == target: normalize(0) is NaN
        c = dict(base, **bad)
        s = Scene(c, None, np.zeros(0, abi.OBJECT_DTYPE)).as_rd_camera()
        before = bytes(s)
        if "target" in bad:  # NaN view: cgmath's unwrap of a NaN determinant
            ra.lib().rd_camera_update_matrices(C.byref(s))
            continue
        assert ra.lib().rd_camera_update_matrices(C.byref(s)) == abi.RD_ERR_INVALID_ARGUMENT, bad
        assert bytes(s) == before


def test_cli_default_scene_is_scene_default(tmp_path):
    """raydar_hip without a scene file renders Scene::default()
    (src/scene/mod.rs:20-67): the scene it writes with --write-scene equals
    scenes/default.rscn, camera matrices bit for bit."""
    import subprocess
    cli = os.path.join(ROOT, "raydar_amd", "raydar_hip")
    out = tmp_path / "default_cli.rscn"
    subprocess.run([cli, "--write-scene", str(out)], check=True, timeout=60)
    got, ref = Scene.load(str(out)), Scene.load(os.path.join(GOLDEN, "default.rscn"))
    for k, v in ref.camera.items():
        g = np.array(got.camera[k], np.float32)
        r = np.array(v, np.float32)
        np.testing.assert_array_equal(g.view(np.uint32), r.view(np.uint32), err_msg=k)
    assert got.world == ref.world
    # the file was saved from an edited scene: its lamp albedo is 0.5, where
    # Scene::default() gives the lamp Material::default()'s 0.8 (material.rs:15-27)
    lamp_ref = ref.objects.copy()
    lamp_ref[2]["albedo"] = np.float32(0.8)
    assert got.objects.tobytes() == lamp_ref.tobytes()


# ------------------------------------------------------ Rust FFI binding ---
RUST_SYS = os.path.join(ROOT, "raydar_amd", "rust", "raydar-hip-sys", "src", "lib.rs")


def _rust_structs(text):
    out = {}
    for name, body in re.findall(r"pub struct (rd_[a-z_]+) \{(.*?)\n\}", text, flags=re.S):
        out[name] = re.findall(r"pub ([a-z_0-9]+): ([^,]+),", body)
    return out


def test_rust_bindings_match_the_c_abi():
    """The FFI crate's #[repr(C)] structs have the header's fields in order
    with the same sizes, and every function it declares is exported by the
    library and declared in include/*.h (no cargo here to compile it)."""
    text = open(RUST_SYS).read()
    structs = _rust_structs(text)
    sizes = {"u32": 4, "i64": 8, "f32": 4, "u64": 8, "*const rd_object": 8}
    mirror = {"rd_object": abi.RdObject, "rd_camera": abi.RdCamera, "rd_world": abi.RdWorld,
              "rd_config": abi.RdConfig, "rd_timings": abi.RdTimings, "rd_scene": abi.RdScene}
    for name, cls in mirror.items():
        fields = structs[name]
        c_fields = cls._fields_
        assert [f for f, _ in fields] == [f for f, _ in c_fields], name
        for (f, rt), (_, ct) in zip(fields, c_fields):
            m = re.fullmatch(r"\[(\w+); (\d+)\]", rt.strip())
            if m:
                n = sizes[m.group(1)] * int(m.group(2))
            elif rt.strip() in sizes:
                n = sizes[rt.strip()]
            else:
                n = C.sizeof(ct)  # nested struct: compared field by field above
            assert n == C.sizeof(ct), (name, f)
    fns = re.findall(r"pub fn (rd_[a-z0-9_]+)\(", text)
    assert len(fns) >= 25
    L = C.CDLL(ra.LIB_PATH)
    declared = set(declared_symbols())
    for f in fns:
        assert hasattr(L, f) and f in declared, f
    hip_rs = open(os.path.join(ROOT, "raydar_amd", "rust", "renderer", "hip.rs")).read()
    for f in re.findall(r"sys::(rd_[a-z0-9_]+)\(", hip_rs):
        assert f in fns, f
