"""Open-loop load driver for the render server.

Requests arrive as a Poisson stream (scenegen.serve_arrivals); each is
submitted at its arrival time whether or not earlier ones have finished, so
requests that arrive while others are on the GPU are admitted as HBM and
queue space free up and their passes run mixed with the others'.  The
server measures each request's time from submit to final image; this module
reports p50/p99 of that, the aggregate path throughput, and the backlog.
"""
import threading
import time

import numpy as np

from .batch import Job, RenderServer


def run(arrivals, hbm_budget_bytes=200 << 30, pass_samples=8, keep_images=False, **server_opts):
    """arrivals: list of (t_seconds, scenegen.Job).  Returns a report dict."""
    # the final RGBA8 image only (no accumulator readback)
    jobs = [Job(j.scene, j.spp, j.bounces, j.seed, pass_samples=pass_samples, accumulator=False)
            for _, j in arrivals]
    results = [None] * len(jobs)
    ids = [None] * len(jobs)
    with RenderServer(hbm_budget_bytes=hbm_budget_bytes, pass_samples=pass_samples,
                      **server_opts) as srv:
        submitted = threading.Event()

        def waiter():
            submitted.wait()
            for k, jid in enumerate(ids):
                results[k] = srv.wait(jid, -1)
                if not keep_images:
                    jobs[k].rgba = None  # release the host image

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        late = 0.0
        for k, (t, _) in enumerate(arrivals):
            now = time.perf_counter() - t0
            if t > now:
                time.sleep(t - now)
            else:
                late = max(late, now - t)
            ids[k] = srv.submit(jobs[k])
        t_submit_end = time.perf_counter() - t0
        submitted.set()
        th.join()
        t_end = time.perf_counter() - t0
        stats = srv.stats()
    lat = np.array([r["latency_ns"] for r in results], np.float64) / 1e6
    queued = np.array([r["queued_ns"] for r in results], np.float64) / 1e6
    paths = sum(j.paths for j in jobs)
    duration = arrivals[-1][0] if arrivals else 0.0
    return {
        "requests": len(jobs),
        "rate_per_s": round(len(jobs) / duration, 3) if duration else None,
        "arrival_span_s": round(duration, 3),
        "offered_mpaths_per_s": round(paths / duration / 1e6, 1) if duration else None,
        "wall_s": round(t_end, 3),
        "drain_after_last_arrival_s": round(t_end - t_submit_end, 3),
        "max_submit_lag_s": round(late, 4),
        "p50_ms": round(float(np.percentile(lat, 50)), 2),
        "p90_ms": round(float(np.percentile(lat, 90)), 2),
        "p99_ms": round(float(np.percentile(lat, 99)), 2),
        "max_ms": round(float(lat.max()), 2),
        "mean_queue_wait_ms": round(float(queued.mean()), 3),
        "aggregate_mpaths_per_s": round(paths / t_end / 1e6, 1),
        "gpu_busy_s": round(stats["busy_ns"] / 1e9, 3),
        "busy_mpaths_per_s": round(paths / max(stats["busy_ns"], 1) * 1e3, 1),
        "hbm_peak_gb": round(stats["hbm_peak"] / 2**30, 3),
        "engine_gb": round(stats["engine_bytes"] / 2**30, 3),
        "max_resident_jobs": stats["max_resident"],
        "all_ok": all(r["status"] == 0 and r["done"] == 1 for r in results),
    }
