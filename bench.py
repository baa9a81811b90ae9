"""Benchmark of the HIP path tracer on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload plain_1080p]

One step renders one whole frame of the workload through the C ABI
(restart_frame + one render pass of all samples + tonemap + readback of the
RGBA8 image), with the scene already resident on the device.  The default
workload is the plain drop-in at production size: 1920x1080, 64 spp, 12
bounces, ground sphere + 64 spheres + 32 cubes (raydar_amd.scenegen
plain_1080p, seed 1).  A path is one camera sample of one pixel; a ray is one
closest-hit query, including the final miss.

Besides the timed steps the script makes three untimed passes over the same
frame: one with algorithmic event counters (FLOPs per the table below), one
with hipEvent timing of every kernel launch (for the extend kernel's roofline
figure), and the CPU baseline (the oracle's restatement of raydar's CPU
renderer on a bounded, stratified pixel sample).

Algorithmic FLOPs per event (1 per IEEE add/sub/mul/div/sqrt/min/max/compare
as written in the reference): camera ray 77 (cpu.rs:199-202,234-251); sphere
test 47 (cpu.rs:39-60); cube test 34 (cpu.rs:65-97); BVH child-box slab test
25; closest_hit sphere 25 / cube 22 (cpu.rs:360-385); scatter non-glass 96 /
glass 215 (cpu.rs:262-333, utils/mod.rs:15-43); sky miss 36 (world.rs:23-27,
cpu.rs:336); accumulate 4 (cpu.rs:204-209); tonemap 16 per pixel.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP = dict(camera=77, sphere_test=47, cube_test=34, box_test=25, hit_sphere=25, hit_cube=22,
            scatter=96, scatter_glass=215, miss=36, accumulate=4, tonemap=16)
PEAK_FP32_TFLOPS = 157.3      # MI355X FP32 vector (packed FMA), AMD public figure
ISSUE_TOPS = 39.3             # 256 CU x 64 lanes x 2.4 GHz, one op per lane per clock
PEAK_HBM_GBS = 8000.0
METRIC = "Mpaths/s at 1080p, 12 bounces (batched vs plain, 1 GPU); p50/p99 render latency"


def extend_flops(c):
    return c["box_tests"] * FLOP["box_test"] + c["sphere_tests"] * FLOP["sphere_test"] + \
        c["cube_tests"] * FLOP["cube_test"]


def path_flops(c):
    hits = c["sphere_hits"] + c["cube_hits"]
    return (c["camera_rays"] * FLOP["camera"] + extend_flops(c) +
            c["sphere_hits"] * FLOP["hit_sphere"] + c["cube_hits"] * FLOP["hit_cube"] +
            (hits - c["glass_scatters"]) * FLOP["scatter"] + c["glass_scatters"] * FLOP["scatter_glass"] +
            c["misses"] * FLOP["miss"] + c["accumulates"] * FLOP["accumulate"] +
            c["tonemapped_px"] * FLOP["tonemap"])


def host_cores():
    """(threads used for the all-core figure, os.cpu_count(), CPUs this process
    may run on).  The GPU box gives a command a 16-core share of its host."""
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return max(1, min(16, allowed)), os.cpu_count(), allowed


def cpu_baseline(job, budget_s=10.0):
    """The oracle's restatement of raydar's CPU renderer on a stratified pixel
    sample of the same frame, 1 thread and all host cores (capped at 16)."""
    import oracle
    sc, spp, bounces = job.scene, job.spp, job.bounces
    npix = sc.width * sc.height
    threads, nproc, allowed = host_cores()
    out = {}
    for nthr in (1, threads):
        stride = 4096 if nthr == 1 else 512
        paths_done, t_used = 0, 0.0
        while True:
            pix = np.arange(0, npix, stride, dtype=np.uint32)
            t0 = time.perf_counter()
            oracle.render(sc, spp, bounces, job.seed, pixels=pix, threads=nthr, fast=True)
            dt = time.perf_counter() - t0
            paths_done, t_used = len(pix) * spp, dt
            if dt >= budget_s / 4 or stride <= 8:
                break
            stride = max(8, int(stride * max(dt, 1e-3) / (budget_s / 2)))
        out[nthr] = (paths_done / t_used / 1e6, paths_done, t_used, stride)
    mt = out[threads]
    return {
        "value": round(mt[0], 4), "unit": "Mpaths/s", "cores": threads, "kind": "port",
        "sample": f"oracle (C++ restatement of raydar's CPU renderer, -O3) on every {mt[3]}th pixel "
                  f"of the same frame x {spp} spp = {mt[1]} paths in {mt[2]:.2f} s on {threads} threads",
        "value_1thread": round(out[1][0], 4),
        "sample_1thread": f"every {out[1][3]}th pixel x {spp} spp = {out[1][1]} paths in {out[1][2]:.2f} s",
        "host_cpu": _cpu_model(),
        "nproc": nproc,
        "cpus_allowed": allowed,
        "build": "g++ -O3 -ffp-contract=off, x86-64 baseline ISA (built in the container, not for "
                 "this host: no -march=native), OpenMP over pixels",
    }


def cpu_baseline_batched(specs, pixels_per_scene=2048):
    """The oracle on the 64-scene batch: every scene's stratified sample of
    pixels_per_scene pixels at its own spp and bounce limit, on the host
    threads.  The rate is the whole batch's paths over the time the sampled
    rates imply for each scene (scenes differ in cost per path: the CPU
    renderer scans every object, cpu.rs:345-349)."""
    import oracle
    threads, nproc, allowed = host_cores()
    est_s, sampled, t_sampled = 0.0, 0, 0.0
    for j in specs:
        npix = j.scene.width * j.scene.height
        k = min(pixels_per_scene, npix)
        step = npix // k
        pix = (np.arange(k, dtype=np.uint64) * step).astype(np.uint32)
        t0 = time.perf_counter()
        oracle.render(j.scene, j.spp, j.bounces, j.seed, pixels=pix, threads=threads, fast=True)
        dt = time.perf_counter() - t0
        est_s += dt / (k * j.spp) * j.paths
        sampled += k * j.spp
        t_sampled += dt
    paths = sum(j.paths for j in specs)
    return {
        "value": round(paths / est_s / 1e6, 4), "unit": "Mpaths/s", "cores": threads, "kind": "port",
        "sample": f"oracle on {pixels_per_scene} stratified pixels of each of the {len(specs)} scenes "
                  f"at its spp and bounce limit: {sampled} paths in {t_sampled:.2f} s on {threads} "
                  f"threads; whole-batch estimate {est_s:.1f} s for {paths} paths",
        "host_cpu": _cpu_model(), "nproc": nproc, "cpus_allowed": allowed,
    }


def batched_vs_plain(reps=2, gap=0.0, cpu=False, opts=None):
    """64 clustered scenes (16-2048 primitives, 128^2..1080p, 8/32/128 spp,
    4/8/12 bounces) in one batched wavefront vs 64 sequential plain renders
    on one reused per-scene renderer (new_frame + render_frame each); each
    side warmed by one untimed pass, then timed over `reps` passes.  `gap`:
    seconds of idle GPU before every pass, so that a kernel trace of the run
    splits into its passes.  `opts`: options set on both renderers."""
    opts = opts or {}
    import raydar_amd as ra
    from raydar_amd import scenegen
    from raydar_amd.batch import BatchRenderer, Job
    specs = scenegen.batch_jobs()
    paths = sum(j.paths for j in specs)
    out = {"jobs": len(specs), "paths": paths}
    with BatchRenderer() as br:
        for k, v in opts.items():
            br.set_option(k, v)
        # one untimed pass of the whole batch (launch shapes, graphs, the
        # kept job buffers), then the mean of `reps` timed passes
        time.sleep(gap)
        br.render([Job(j.scene, j.spp, j.bounces, j.seed, accumulator=False) for j in specs])
        runs = []
        for _ in range(reps):
            jobs = [Job(j.scene, j.spp, j.bounces, j.seed, accumulator=False) for j in specs]
            time.sleep(gap)
            t0 = time.perf_counter()
            st = br.render(jobs)
            runs.append(time.perf_counter() - t0)
        dt_b = sum(runs) / len(runs)
        out["batched"] = {"seconds": round(dt_b, 4), "mpaths_per_s": round(paths / dt_b / 1e6, 2),
                          "runs_s": [round(t, 4) for t in runs],
                          "iterations": st["iterations"], "graph_launches": st["graph_launches"],
                          "options": {k: br.get_option(k) for k in ("block", "queue", "leaf", "pipes",
                                                                    "solo_paths")},
                          "solo_jobs": br.get_option("solo_jobs")}
    solo_paths = out["batched"]["options"]["solo_paths"]
    # the solo-lane threshold is the batch renderer's own option
    with ra.HipRenderer(**{k: v for k, v in opts.items() if k != "solo_paths"}) as r:
        runs = []
        large_s = [0.0] * (reps + 1)  # time of the jobs the batch gives the solo lane
        for k in range(reps + 1):  # the first pass warms up, as for the batch
            time.sleep(gap)
            t0 = time.perf_counter()
            for j in specs:
                tj = time.perf_counter()
                r.max_sample_count = j.spp
                r.max_bounces = j.bounces
                r.set_seed(j.seed)
                r.render_frame(j.scene)
                if solo_paths and j.paths >= solo_paths:
                    large_s[k] += time.perf_counter() - tj
            if k:
                runs.append(time.perf_counter() - t0)
        dt_p = sum(runs) / len(runs)
        dt_large = sum(large_s[1:]) / reps
    n_large = sum(1 for j in specs if solo_paths and j.paths >= solo_paths)
    out["sequential_plain"] = {"seconds": round(dt_p, 4), "mpaths_per_s": round(paths / dt_p / 1e6, 2),
                               "runs_s": [round(t, 4) for t in runs],
                               # the batch can beat the sequential renders by at most this
                               # much if its small jobs cost nothing and its large ones run
                               # as fast as plain renders
                               "large_jobs": n_large,
                               "large_job_paths": sum(j.paths for j in specs
                                                      if solo_paths and j.paths >= solo_paths),
                               "large_jobs_s": round(dt_large, 4),
                               "bound_if_small_jobs_were_free": round(dt_p / dt_large, 3)
                               if dt_large > 0 else None}
    if dt_large > 0:
        # the whole batch at the large jobs' plain rate (big passes, no
        # per-frame overheads): if the small jobs' paths cost as much as the
        # large jobs' (the same scene generator, spp and bounce choices), a
        # batch at that rate beats the sequential renders by this much
        large_rate = out["sequential_plain"]["large_job_paths"] / dt_large
        out["sequential_plain"]["large_jobs_mpaths_per_s"] = round(large_rate / 1e6, 1)
        out["sequential_plain"]["bound_at_large_job_rate"] = round(dt_p / (paths / large_rate), 3)
    out["batched_over_plain"] = round(dt_p / dt_b, 3)
    if cpu:
        print("# cpu baseline on the batch...", file=sys.stderr, flush=True)
        cb = cpu_baseline_batched(specs)
        out["cpu_baseline_batched"] = cb
        out["batched_vs_cpu"] = round(out["batched"]["mpaths_per_s"] / cb["value"], 1)
    return out


def serving(seconds=60.0, rate=20.0):
    """Poisson request stream against the render server (200 GB HBM budget,
    8-spp progressive passes, scenes as the batch with 256^2..1080p,
    32/64/128 spp, 12 bounces); latency = submit to final image."""
    from raydar_amd import scenegen, serving as srv
    arrivals = scenegen.serve_arrivals(rate=rate, duration=seconds)
    rep = srv.run(arrivals, hbm_budget_bytes=200 << 30, pass_samples=8)
    rep["config"] = f"Poisson {rate}/s for {seconds} s, arrival seed 7, scene seeds 2000+i"
    return rep


def calibrate_copy(nbytes):
    """The library's plain copy kernel (k_copy_calibration: read nbytes, write
    nbytes) five times: what FETCH_SIZE / WRITE_SIZE read for a known amount
    of HBM traffic."""
    import raydar_amd as ra
    ms = ra.debug_copy_calibration(nbytes, 5)
    print(json.dumps({"calibrate_copy_bytes": nbytes, "launches": 5,
                      "ms": [round(float(x), 4) for x in ms],
                      "GB_per_s": round(2 * nbytes / (float(np.median(ms)) * 1e-3) / 1e9, 1)}),
          flush=True)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "extend_plain_1080p_hbm_traffic.md")


KERNEL_SOURCES = ("kernels.hip", "pathmath.hpp", "layout.hpp", "engine.hip")


def kernel_sources_sha16():
    """Hash of the sources that decide the extend kernel's memory traffic."""
    import hashlib
    h = hashlib.sha256()
    for name in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "raydar_amd", "csrc", name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def extend_traffic(workload):
    """HBM bytes per extend launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by the copy-kernel calibration).
    The counters cannot be read from inside this process, so the figure is
    the one measured for this workload by the command the summary names, and
    only while the kernel sources are the ones it was measured on (the
    summary records their hash): (bytes or None, note)."""
    if workload != "plain_1080p" or not os.path.exists(TRAFFIC_FILE):
        return None, "no counter summary for this workload"
    vals = {}
    with open(TRAFFIC_FILE) as f:
        for line in f:
            if line.startswith("traffic."):
                k, v = line[len("traffic."):].split("=", 1)
                vals[k.strip()] = v.strip()
    if "bytes_per_launch" not in vals:
        return None, "counter summary without traffic.bytes_per_launch"
    now = kernel_sources_sha16()
    if vals.get("kernel_sources_sha16") != now:
        return None, (f"counter summary measured on kernel sources {vals.get('kernel_sources_sha16')}, "
                      f"current {now}: stale, re-run its commands")
    return round(float(vals["bytes_per_launch"])), \
        f"rocprofv3 FETCH_SIZE + WRITE_SIZE of these kernel sources ({now}), {os.path.relpath(TRAFFIC_FILE, ROOT)}"


def _cpu_model():
    """Model name plus family / model / stepping and clock from /proc/cpuinfo
    (virtual machines often report a bare 'AMD EPYC' model name)."""
    fields = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    if fields:
                        break  # the first processor's block is enough
                    continue
                k, v = (s.strip() for s in line.split(":", 1))
                fields.setdefault(k, v)
    except OSError:
        return "unknown"
    name = fields.get("model name", "unknown")
    extra = [f"{k} {fields[k]}" for k in ("cpu family", "model", "stepping") if k in fields]
    if "cpu MHz" in fields:
        extra.append(f"{float(fields['cpu MHz']):.0f} MHz")
    return f"{name} ({', '.join(extra)})" if extra else name


def extend_simd(c):
    """SIMD efficiency of the closest-hit kernel from the counting pass: the
    share of a wave's 64 lanes that hold a ray per traversal iteration, and
    per wave iteration the share that orders entered children, the primitive
    loop's trips and the share of those trips that test a cube."""
    ws = max(c["wave_steps"], 1)
    trips = max(c["prim_trips"], 1)
    return {
        "lane_utilization": round(c["lane_steps"] / (64.0 * ws), 4),
        "sort_share": round(c["sort_steps"] / ws, 4),
        "prim_trips_per_step": round(c["prim_trips"] / ws, 4),
        "prim_lanes_per_trip": round((c["sphere_tests"] + c["cube_tests"]) / trips, 2),
        "cube_trip_share": round(c["cube_trips"] / trips, 4),
        "box_tests_per_lane_step": round(c["box_tests"] / max(c["lane_steps"], 1), 3),
        "steps_per_refill": round(ws / max(c["refills"], 1), 2),
        "wave_steps": c["wave_steps"],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="plain_1080p")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the batched-vs-plain comparison (for runs under a profiler)")
    ap.add_argument("--phase-gap", type=float, default=0.0,
                    help="seconds of idle GPU before each pass of the batched comparison")
    ap.add_argument("--autotune", action="store_true",
                    help="refused: a full run with this autotune and a 30 s server once ended "
                         "in a GPU fault whose cause has not been shown")
    ap.add_argument("--serve-seconds", type=float, default=1.5,
                    help="length of the Poisson serving run (default 1.5 s, the length the "
                         "GPU tests run; 0: skipped; longer runs are refused, as --autotune is)")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--option", action="append", default=[], help="key=value renderer option")
    ap.add_argument("--timed-only", action="store_true",
                    help="only the warmup and timed steps (for runs under a profiler)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="run a plain device copy of BYTES bytes 5 times and exit "
                         "(the reference point for HBM counter readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(this build renders on a single MI355X)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    # A run of the full benchmark with a 30 s server and the 8-scene autotune
    # once ended in a GPU fault, and no cause has been shown (host
    # orchestration cleared on a fake device under sanitizers): those two
    # cases stay off the GPU.
    if args.serve_seconds > 1.5 or args.autotune:
        print("bench.py: serving runs longer than 1.5 s and --autotune are refused until the cause "
              "of an earlier GPU fault (a full run with a 30 s server and the autotune) is shown",
              file=sys.stderr)
        sys.exit(2)

    if args.batched_only:
        opts = {k: int(v) for k, v in (kv.split("=", 1) for kv in args.option)}
        print(json.dumps({"batched": batched_vs_plain(gap=args.phase_gap, opts=opts)}),
              flush=True)
        return

    import raydar_amd as ra
    from raydar_amd import scenegen

    job = scenegen.WORKLOADS[args.workload]()
    sc = job.scene
    paths_per_step = sc.width * sc.height * job.spp
    opts = dict(kv.split("=", 1) for kv in args.option)
    r = ra.HipRenderer(max_sample_count=job.spp, max_bounces=job.bounces, seed=job.seed,
                       **{k: int(v) for k, v in opts.items()})
    r.set_samples_per_pass(job.spp)
    r.new_frame(sc)

    frame = np.zeros((sc.height, sc.width, 4), np.uint8)  # the caller's frame buffer, reused

    def step():
        r.restart_frame()
        img = r.render_sample(out=frame)
        assert img is not None and r.sample_count == job.spp

    for _ in range(args.warmup):
        step()
    print(f"# warmup done ({args.warmup} steps)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1e3
    value = paths_per_step * args.steps / elapsed / 1e6
    print(f"# timed {args.steps} steps: {ms_per_step:.2f} ms/step, {value:.1f} Mpaths/s",
          file=sys.stderr, flush=True)
    if args.timed_only:
        print(json.dumps({"metric": METRIC, "value": round(value, 3), "unit": "Mpaths/s",
                          "ms_per_step": round(ms_per_step, 3), "steps": args.steps,
                          "warmup": args.warmup, "timed_only": True}), flush=True)
        r.close()
        return

    # counting pass (algorithmic events)
    r.reset_counters()
    r.set_option("count", 1)
    step()
    r.set_option("count", 0)
    cnt = r.counters()
    # kernel-timing pass: hipEvents around every launch on the stream it runs
    # on, each launch waited for before the next, so every kernel is timed
    # alone (in the timed steps the pipes' kernels overlap)
    pipes_used = r.get_option("pipes_used")
    r.reset_counters()
    r.set_option("kernel_timing", 1)
    step()
    r.set_option("kernel_timing", 0)
    kt = r.kernel_times()
    iters = r.get_option("iterations")
    ext_ms, ext_n = kt["extend"]
    ext_flop_per_launch = extend_flops(cnt) / max(ext_n, 1)
    ext_avg_s = ext_ms / 1e3 / max(ext_n, 1)
    achieved = ext_flop_per_launch / ext_avg_s / 1e12 if ext_avg_s > 0 else 0.0
    # the same kernel on every camera ray (hit_reuse 0: each sample traces
    # its pixel's camera ray, as the reference does): one counting and one
    # kernel-timing frame, for comparison with the production figure above
    r.set_option("hit_reuse", 0)
    r.reset_counters()
    r.set_option("count", 1)
    step()
    r.set_option("count", 0)
    cnt_nr = r.counters()
    r.reset_counters()
    r.set_option("kernel_timing", 1)
    step()
    r.set_option("kernel_timing", 0)
    kt_nr = r.kernel_times()
    r.set_option("hit_reuse", 1)
    nr_ms, nr_n = kt_nr["extend"]
    nr_flop = extend_flops(cnt_nr) / max(nr_n, 1)
    nr_s = nr_ms / 1e3 / max(nr_n, 1)
    nr_achieved = nr_flop / nr_s / 1e12 if nr_s > 0 else 0.0
    rays_per_step = cnt["segments"]
    traffic, traffic_note = extend_traffic(args.workload)
    total_flop = path_flops(cnt)
    step_s = ms_per_step / 1e3

    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {sc.width}x{sc.height}, {job.spp} spp, {job.bounces} bounces, "
                        f"{len(sc.objects)} objects, seed {job.seed}",
            "paths_per_step": paths_per_step,
            "renderer": "plain per-scene HIP renderer (wavefront: generate/extend/shade, SAH BVH)",
            "options": {k: r.get_option(k) for k in ("block", "queue", "leaf", "grid", "traversal",
                                                      "bvh4", "pipes")},
            "pipes_used": pipes_used,
        },
        "mrays_per_s": round(rays_per_step / step_s / 1e6, 3),
        "rays_per_path": round(rays_per_step / paths_per_step, 4),
        "extend_simd": extend_simd(cnt),
        "whole_path_tflops": round(total_flop / step_s / 1e12, 4),
        "whole_path_frac_fp32_peak": round(total_flop / step_s / 1e12 / PEAK_FP32_TFLOPS, 5),
        "roofline": {
            "bound": "fp32_valu",
            "compute": "FP32 vector ALU (this path has no dense contraction and uses no MFMA); "
                       "peak is MI355X FP32 157.3 TFLOP/s",
            "kernel": "extend (closest hit: BVH traversal + exact sphere/cube tests)",
            "achieved": round(achieved, 4),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 6),
            "frac_of_issue_ceiling": round(achieved / ISSUE_TOPS, 6),
            "traffic": traffic,
            "traffic_source": traffic_note,
            "flop_per_launch": round(ext_flop_per_launch),
            "avg_launch_ms": round(ext_avg_s * 1e3, 5),
            "launches": ext_n,
            "timing": "each launch timed alone (kernel-timing pass, launches serialized)",
            "every_camera_ray_traced": {
                "note": "same kernel with option hit_reuse 0 (each sample traces its pixel's "
                        "camera ray, as the reference does; same image): the camera rays it "
                        "then traces are its cheapest, most coherent work",
                "achieved": round(nr_achieved, 4),
                "frac": round(nr_achieved / PEAK_FP32_TFLOPS, 6),
                "frac_of_issue_ceiling": round(nr_achieved / ISSUE_TOPS, 6),
                "avg_launch_ms": round(nr_s * 1e3, 5),
                "launches": nr_n,
                "rays_per_step": cnt_nr["segments"],
                "extend_ms_per_step": round(nr_ms, 3),
            },
        },
        "kernel_ms_per_step": {k: round(v[0], 3) for k, v in kt.items()},
        "kernel_launches_per_step": {k: v[1] for k, v in kt.items()},
        "iterations_per_step": iters,
        "counters_per_step": cnt,
    }
    r.close()
    if not args.no_batched:
        print("# batched vs plain...", file=sys.stderr, flush=True)
        result["batched"] = batched_vs_plain(cpu=not args.no_cpu_baseline)
    if args.serve_seconds > 0:
        print(f"# serving {args.serve_seconds} s...", file=sys.stderr, flush=True)
        result["serving"] = serving(args.serve_seconds, args.serve_rate)
    if not args.no_cpu_baseline:
        print("# cpu baseline...", file=sys.stderr, flush=True)
        cb = cpu_baseline(job, args.cpu_budget)
        result["cpu_baseline"] = cb
        result["gpu_vs_cpu"] = round(value / cb["value"], 2)
        result["gpu_vs_cpu_1thread"] = round(value / cb["value_1thread"], 2)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
