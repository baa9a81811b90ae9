"""The host-side scene reader, BVH builder, traversals and radiance ring under
AddressSanitizer and UndefinedBehaviorSanitizer (tests/host_check.cpp; CPU
only, built with g++ from the library's host sources)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_host_code_under_sanitizers(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "host_check")
    src = [os.path.join(ROOT, "tests", "host_check.cpp"),
           os.path.join(ROOT, "raydar_amd", "csrc", "bvh.cpp"),
           os.path.join(ROOT, "raydar_amd", "csrc", "rscn.cpp")]
    flags = ["-std=c++17", "-O1", "-g", "-ffp-contract=off", "-fno-omit-frame-pointer",
             "-fsanitize=address,undefined,float-cast-overflow",
             "-fno-sanitize-recover=undefined,float-cast-overflow",
             "-I", os.path.join(ROOT, "include")]
    build = subprocess.run(["g++", *flags, *src, "-o", exe], capture_output=True, text=True)
    if build.returncode != 0 and "asan" in build.stderr.lower() and "cannot find" in build.stderr:
        pytest.skip("sanitizer runtime not installed")
    assert build.returncode == 0, build.stderr[-4000:]
    scenes = [os.path.join(GOLDEN, n) for n in ("default.rscn", "benchmark.rscn")]
    # verify_asan_link_order=0: the process may preload other libraries
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([exe, *scenes], capture_output=True, text=True, timeout=600, env=env)
    assert run.returncode == 0, (run.stdout[-2000:], run.stderr[-6000:])
    assert "host checks ok" in run.stdout
