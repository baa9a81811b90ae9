"""Batched renderer and render server on the GPU.

Every job's image depends only on (scene, seed, samples, bounces): the tests
render jobs alone, mixed with others, split into passes, through the
captured graph and through direct launches, and require bit-identical
accumulators, equal to the oracle's FP32 render of the same job."""
import os
import time

import numpy as np
import pytest

import oracle
import raydar_amd as ra
from raydar_amd import abi, scenegen
from raydar_amd.batch import BatchRenderer, Job, RenderServer
from raydar_amd.scene import Scene, make_camera, sky_world

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _jobs():
    empty = Scene(make_camera((0, 1, 5), (0, 0, 0), (0, 1, 0), 40, 30, fov_deg=60), sky_world(),
                  np.zeros(0, abi.OBJECT_DTYPE))
    return [
        (scenegen.plain_small(0), 4, 6, 11),
        (scenegen.clustered(5, 300, 96, 64), 3, 12, 12),
        (scenegen.plain_1080p(1, 120, 68), 2, 12, 13),
        (Scene.load(os.path.join(GOLDEN, "default.rscn")), 1, 12, 14),
        (empty, 2, 4, 15),
    ]


def _oracle_acc(scene, spp, bounces, seed):
    return oracle.render(scene, spp, bounces, seed)


def test_batch_matches_oracle_and_plain():
    specs = _jobs()
    jobs = [Job(s, spp, b, seed) for s, spp, b, seed in specs]
    with BatchRenderer() as br:
        st = br.render(jobs)
    assert st["paths"] == sum(j.paths for j in jobs)
    for job, (s, spp, b, seed) in zip(jobs, specs):
        ref = _oracle_acc(s, spp, b, seed)
        np.testing.assert_array_equal(job.acc.view(np.uint32), ref.view(np.uint32))
        np.testing.assert_array_equal(job.rgba.reshape(-1), oracle.tonemap(ref, spp))
        with ra.HipRenderer(max_sample_count=spp, max_bounces=b, seed=seed) as r:
            r.render_frame(s)
            np.testing.assert_array_equal(r.accumulator(), job.acc)


@pytest.mark.parametrize("opts", [dict(graph=0), dict(graph=1, block=128, grid=97),
                                  dict(block=64, leaf=1), dict(queue=4096), dict(refill=0),
                                  dict(refill=1, queue=4096), dict(pipes=1), dict(pipes=2, queue=4096),
                                  dict(pipes=4, graph=0), dict(bvh4=0), dict(bvh4=0, graph=0, block=128)])
def test_batch_launch_shapes_and_passes_agree(opts):
    specs = _jobs()[:3]
    a = [Job(s, spp, b, seed) for s, spp, b, seed in specs]
    b_ = [Job(s, spp, b, seed, pass_samples=1) for s, spp, b, seed in specs]
    with BatchRenderer(**opts) as br:
        br.render(a)
    with BatchRenderer() as br:
        br.render(b_)
    for x, y in zip(a, b_):
        np.testing.assert_array_equal(x.acc, y.acc)
        np.testing.assert_array_equal(x.rgba, y.rgba)


def test_batch_pipes_option():
    specs = _jobs()
    jobs = [Job(s, spp, b, seed) for s, spp, b, seed in specs]
    with BatchRenderer() as br:
        assert br.get_option("pipes") == 3
        br.set_option("pipes", 2)
        assert br.get_option("pipes") == 2
        for bad in (0, 5):
            with pytest.raises(ra.RaydarError) as e:
                br.set_option("pipes", bad)
            assert e.value.status == abi.RD_ERR_INVALID_ARGUMENT
        br.render(jobs)
        with pytest.raises(ra.RaydarError) as e:  # engines allocated: fixed from now on
            br.set_option("pipes", 3)
        assert e.value.status == abi.RD_ERR_INVALID_ARGUMENT
        br.set_option("pipes", 2)
    for job, (s, spp, b, seed) in zip(jobs, specs):
        np.testing.assert_array_equal(job.acc.view(np.uint32), _oracle_acc(s, spp, b, seed).view(np.uint32))


@pytest.mark.parametrize("pipes", [1, 3])
def test_batch_ticket_list_rewinds_at_cap(pipes):
    # a 2-entry ticket table: passes wait for the work in flight to drain,
    # then the list starts over; results are unchanged
    specs = _jobs()[:3]
    jobs = [Job(s, spp, b, seed, pass_samples=1) for s, spp, b, seed in specs]
    with BatchRenderer(pipes=pipes, ticket_cap=2) as br:
        assert br.get_option("ticket_cap") == 2
        br.render(jobs)
        assert br.get_option("rewinds") > 0
        with pytest.raises(ra.RaydarError):
            br.set_option("ticket_cap", 4)
    for job, (s, spp, b, seed) in zip(jobs, specs):
        np.testing.assert_array_equal(job.acc.view(np.uint32), _oracle_acc(s, spp, b, seed).view(np.uint32))


def test_server_alone_equals_mixed():
    specs = _jobs()
    target = specs[1]
    with RenderServer(pass_samples=1) as srv:
        alone = Job(*target[:1], target[1], target[2], target[3])
        jid = srv.submit(alone)
        r = srv.wait(jid, 60000)
        assert r["done"] == 1 and r["status"] == 0 and r["latency_ns"] > 0
    with RenderServer(pass_samples=2) as srv:
        mixed = [Job(s, spp, b, seed, progress=True) for s, spp, b, seed in specs]
        ids = []
        for j in mixed:
            ids.append(srv.submit(j))
            time.sleep(0.002)  # arrivals while others are on the GPU
        res = [srv.wait(i, 60000) for i in ids]
        st = srv.stats()
    assert all(r["done"] == 1 and r["status"] == 0 for r in res)
    np.testing.assert_array_equal(mixed[1].acc, alone.acc)
    np.testing.assert_array_equal(mixed[1].rgba, alone.rgba)
    for j in mixed:
        np.testing.assert_array_equal(j.progress, j.rgba)  # last progressive image = final
    assert st["jobs_done"] == len(specs) and st["paths"] == sum(j.paths for j in mixed)
    ref = _oracle_acc(*target)
    np.testing.assert_array_equal(alone.acc.view(np.uint32), ref.view(np.uint32))


def test_server_budget_admission():
    sc = scenegen.plain_small(0)
    with pytest.raises(ra.RaydarError) as e:
        RenderServer(hbm_budget_bytes=1 << 20)
    assert e.value.status == abi.RD_ERR_BUDGET
    with RenderServer(queue=65536) as probe:
        fixed = probe.stats()["engine_bytes"]
        jid = probe.submit(Job(sc, 1, 1, 0))
        job_bytes = probe.wait(jid, 60000)["device_bytes"]
    assert fixed > 0 and job_bytes > 256 * 256 * 20
    # room for the engine and one job: the others wait for admission
    with RenderServer(hbm_budget_bytes=fixed + job_bytes * 3 // 2, queue=65536) as srv:
        jobs = [Job(sc, 2, 4, s) for s in range(6)]
        ids = [srv.submit(j) for j in jobs]
        res = [srv.wait(i, 60000) for i in ids]
        st = srv.stats()
    assert all(r["done"] and r["status"] == 0 for r in res)
    assert st["max_resident"] == 1 and st["hbm_peak"] <= job_bytes
    assert max(r["queued_ns"] for r in res) > 0
    for j, s in zip(jobs, range(6)):
        np.testing.assert_array_equal(j.acc, _oracle_acc(sc, 2, 4, s))


def test_server_rejects_invalid_jobs():
    sc = scenegen.plain_small(0)
    with RenderServer(queue=65536) as srv:
        with pytest.raises(ra.RaydarError) as e:
            srv.submit(Job(sc, 0, 4, 0))
        assert e.value.status == abi.RD_ERR_INVALID_ARGUMENT
        t = scenegen.plain_small(0)
        t.world = dict(t.world, kind=abi.RD_WORLD_TRANSPARENT)
        with pytest.raises(ra.RaydarError) as e:
            srv.submit(Job(t, 1, 4, 0))
        assert e.value.status == abi.RD_ERR_UNSUPPORTED
        assert srv.stats()["jobs_rejected"] == 2


def test_poisson_serving_run_completes():
    from raydar_amd import serving
    arrivals = scenegen.serve_arrivals(rate=20.0, duration=1.5)
    rep = serving.run(arrivals, hbm_budget_bytes=200 << 30, pass_samples=8)
    print(rep)
    assert rep["all_ok"] and rep["requests"] == len(arrivals)
    assert 0 < rep["p50_ms"] <= rep["p99_ms"] <= rep["max_ms"]


def test_autotune_returns_a_valid_setting():
    jobs = [Job(scenegen.plain_small(0), 1, 4, 0), Job(scenegen.clustered(5, 200, 64, 64), 1, 4, 1)]
    with BatchRenderer() as br:
        r = br.autotune(jobs)
        assert r["block"] in (64, 128, 256) and r["leaf"] in (1, 2, 4, 8)
        assert r["configs_tried"] == 48 and r["best_ms"] <= r["default_ms"] and not r["cached"]
        again = br.autotune(jobs)  # the same mix on the same GPU: kept, nothing timed
        assert again["cached"] and {k: again[k] for k in ("block", "leaf", "queue")} == \
            {k: r[k] for k in ("block", "leaf", "queue")}
        out = [Job(scenegen.plain_small(0), 2, 4, 0)]
        br.render(out)
    np.testing.assert_array_equal(out[0].acc, _oracle_acc(scenegen.plain_small(0), 2, 4, 0))


def _stratified_pixels(scene, k, salt):
    npix = scene.width * scene.height
    k = min(k, npix)
    step = npix // k
    return (np.arange(k, dtype=np.uint64) * step + (salt * 7919) % max(1, step)).astype(np.uint32)


def test_batch_64_scene_workload_matches_oracle():
    """The full 64-scene batch (seeds 1000-1063: 16-2048 primitives, 128^2 to
    1920x1080, 8/32/128 spp, 4/8/12 bounces; 2.05 G paths) with the default
    pipes and solo lane: every job's accumulator equals the oracle's FP32
    render bit for bit on 2048 stratified pixels, and its image is the
    oracle's tonemap there."""
    jobs_spec = scenegen.batch_jobs()
    jobs = [Job(j.scene, j.spp, j.bounces, j.seed) for j in jobs_spec]
    with BatchRenderer() as br:
        st = br.render(jobs)
        solo = br.get_option("solo_jobs")
    assert st["paths"] == sum(j.paths for j in jobs) == 2047647744
    assert 0 < solo < len(jobs)  # both the shared queues and the solo lane rendered jobs
    threads = max(1, min(16, os.cpu_count() or 1))
    for job, spec in zip(jobs, jobs_spec):
        pix = _stratified_pixels(spec.scene, 2048, spec.seed)
        ref = oracle.render(spec.scene, spec.spp, spec.bounces, spec.seed, pixels=pix, threads=threads)
        got = job.acc[pix]
        np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32),
                                      err_msg=f"job seed {spec.seed}")
        np.testing.assert_array_equal(job.rgba.reshape(-1, 4)[pix].reshape(-1),
                                      oracle.tonemap(ref, spec.spp))


_BIG_REF = {}


@pytest.mark.parametrize("solo_paths", [0, 1 << 21])
def test_batch_full_size_job_among_small_ones(solo_paths):
    # a full 1080p job (2 spp, 4.1 M paths) sharing three pipes with small
    # jobs, or (solo_paths 2^21) rendered by the solo lane beside them; it
    # must come out as rendered alone, equal to the oracle
    big = (scenegen.plain_1080p(1), 2, 12, 21)
    specs = [big] + _jobs()[:3]
    jobs = [Job(s, spp, b, seed) for s, spp, b, seed in specs]
    with BatchRenderer(solo_paths=solo_paths) as br:
        br.render(jobs)
        assert br.get_option("solo_jobs") == (1 if solo_paths else 0)
    if "big" not in _BIG_REF:
        _BIG_REF["big"] = oracle.render(*big, threads=max(1, min(16, os.cpu_count() or 1)))
    ref = _BIG_REF["big"]
    np.testing.assert_array_equal(jobs[0].acc.view(np.uint32), ref.view(np.uint32))
    np.testing.assert_array_equal(jobs[0].rgba.reshape(-1), oracle.tonemap(ref, 2))
    for job, (s, spp, b, seed) in zip(jobs[1:], specs[1:]):
        np.testing.assert_array_equal(job.acc.view(np.uint32), _oracle_acc(s, spp, b, seed).view(np.uint32))


def test_batch_solo_lane_routing_and_results():
    # jobs of at least solo_paths paths (and no progress image) render on the
    # solo lane's per-scene engine, the rest on the shared queues; every
    # image equals the oracle's either way
    specs = _jobs()
    with BatchRenderer() as br:
        assert br.get_option("solo_paths") == 64 << 20
        with pytest.raises(ra.RaydarError) as e:
            br.set_option("solo_paths", -1)
        assert e.value.status == abi.RD_ERR_INVALID_ARGUMENT
        for threshold in (1, 20000):
            br.set_option("solo_paths", threshold)
            jobs = [Job(s, spp, b, seed) for s, spp, b, seed in specs]
            # above the threshold but with a progress image: stays on the queues
            jobs.append(Job(specs[0][0], 2, 6, 31, progress=True))
            before = br.get_option("solo_jobs")
            st = br.render(jobs)
            assert st["paths"] == sum(j.paths for j in jobs)
            lane = sum(1 for j in jobs[:-1] if j.paths >= threshold)
            assert 0 < lane < len(jobs)
            assert br.get_option("solo_jobs") - before == lane
            for job, (s, spp, b, seed) in zip(jobs, specs):
                ref = _oracle_acc(s, spp, b, seed)
                np.testing.assert_array_equal(job.acc.view(np.uint32), ref.view(np.uint32))
                np.testing.assert_array_equal(job.rgba.reshape(-1), oracle.tonemap(ref, spp))
            ref = _oracle_acc(specs[0][0], 2, 6, 31)
            np.testing.assert_array_equal(jobs[-1].acc.view(np.uint32), ref.view(np.uint32))
        # a second call reuses the lane's engine
        br.set_option("solo_paths", 1)
        again = [Job(s, spp, b, seed) for s, spp, b, seed in specs[:2]]
        br.render(again)
    for job, (s, spp, b, seed) in zip(again, specs[:2]):
        np.testing.assert_array_equal(job.acc.view(np.uint32), _oracle_acc(s, spp, b, seed).view(np.uint32))
