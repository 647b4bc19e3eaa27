"""The library's threaded host code for chunked moving averages
(lbm_mv_avg_chunk in lbm_amd/csrc/d2q9_observe.cpp) under ThreadSanitizer:
tests/observe_sanitize_main.cpp computes a chunk on the threaded path and on
one thread and requires equal results; built with g++ -fsanitize=thread and
run as a child process, so that a data race fails it.  CPU only (skipped
where g++ or the sanitizer runtime is missing)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_moving_averages_thread_clean(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no g++")
    exe = tmp_path / "observe_tsan"
    cmd = [cxx, "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-pthread",
           os.path.join(ROOT, "lbm_amd", "csrc", "d2q9_observe.cpp"),
           os.path.join(ROOT, "tests", "observe_sanitize_main.cpp"), "-o", str(exe)]
    build = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    if build.returncode != 0 and "tsan" in (build.stderr + build.stdout).lower():
        pytest.skip("sanitizer runtime not available: " + build.stderr[-300:])
    assert build.returncode == 0, build.stderr
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1:exitcode=66")
    run = subprocess.run([str(exe), "8"], capture_output=True, text=True, timeout=600, env=env)
    assert run.returncode == 0, (run.returncode, run.stdout[-2000:], run.stderr[-4000:])
    assert run.stdout.startswith("ok "), run.stdout
