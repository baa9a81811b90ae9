"""The library's host code on a fake HIP device (tests/fakehip), on the CPU.

engine.hip, batch.hip and api.hip are compiled unchanged with g++ against a
host-only HIP runtime (streams are threads that run their operations in
order after random delays; device memory is host memory in a registry) and
fake kernels that keep the wavefront's bookkeeping and check every pointer,
table index, radiance range, upload order and accumulation order the real
kernels depend on.  The stress driver runs the benchmark's workloads through
it: per-scene frames over 1-4 pipes, the 64-scene batch (with forced ticket
rewinds and radiance-ring wrap-around), the 48-shape autotune, and the
Poisson server stream of the serving workload (20 requests/s for 60 s, time
compressed), under AddressSanitizer + UndefinedBehaviorSanitizer and under
ThreadSanitizer, with hipFree both waiting for queued work and not waiting
(on the MI355X a free does not order itself after work on other streams).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FAKEHIP = os.path.join(ROOT, "tests", "fakehip")


def _build(out, target):
    if shutil.which("g++") is None or shutil.which("make") is None:
        pytest.skip("g++ / make not available")
    r = subprocess.run(["make", "-s", "-j8", "-C", FAKEHIP, f"OUT={out}", target],
                       capture_output=True, text=True, timeout=1200)
    if r.returncode != 0 and "cannot find" in r.stderr and "san" in r.stderr:
        pytest.skip("sanitizer runtime not installed")
    assert r.returncode == 0, r.stderr[-4000:]
    return os.path.join(out, target)


@pytest.fixture(scope="module")
def asan(tmp_path_factory):
    return _build(str(tmp_path_factory.mktemp("fakehip")), "stress_asan")


@pytest.fixture(scope="module")
def tsan(tmp_path_factory):
    return _build(str(tmp_path_factory.mktemp("fakehip_tsan")), "stress_tsan")


def _run(exe, args, timeout, **env):
    e = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
             UBSAN_OPTIONS="print_stacktrace=1", TSAN_OPTIONS="halt_on_error=1", **env)
    r = subprocess.run([exe, *args, "--delay-us", "20"], capture_output=True, text=True,
                       timeout=timeout, env=e)
    assert r.returncode == 0 and "stress ok" in r.stdout, (r.stdout[-3000:], r.stderr[-6000:])
    return r.stdout


@pytest.mark.parametrize("args,nosync", [
    (["plain"], "0"),
    (["noalloc"], "1"),
    (["batch"], "1"),
    (["autotune"], "0"),
    (["server", "--seconds", "60", "--speed", "20"], "1"),
    (["server_budget", "--seconds", "20", "--speed", "20"], "0"),
])
def test_host_code_on_fake_device_asan(asan, args, nosync):
    out = _run(asan, args, 900, FAKEHIP_FREE_NOSYNC=nosync)
    if args[0] == "server":
        assert "1235 requests, 1235 done" in out  # the serving workload's 60 s stream


@pytest.mark.parametrize("args", [["plain"], ["server", "--seconds", "6", "--speed", "10"]])
def test_host_code_on_fake_device_tsan(tsan, args):
    _run(tsan, args, 900)


def test_autotune_results_kept_per_device_across_processes(asan, tmp_path):
    # the first process tunes both mixes and appends them to the file; the
    # second finds both there and times nothing (stress.cpp checks that a
    # repeated call within a process launches no kernel)
    f = str(tmp_path / "tune.txt")
    first = _run(asan, ["autotune"], 900, RAYDAR_HIP_TUNE_CACHE=f)
    assert "autotune cached=0" in first and "other mix cached=0" in first
    lines = open(f).read().splitlines()
    assert len(lines) == 2 and all("gfx950" in l and l.count("\t") == 6 for l in lines)
    second = _run(asan, ["autotune"], 900, RAYDAR_HIP_TUNE_CACHE=f)
    assert "autotune cached=1" in second and "other mix cached=1" in second
    assert len(open(f).read().splitlines()) == 2
