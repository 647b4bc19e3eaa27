"""The C oracle under AddressSanitizer and UndefinedBehaviorSanitizer: the
driver tests/oracle_sanitize_main.c runs the reference's call order on a
channel with interpolated and plain bounce-back links (both IBB branches,
links whose stencils wrap at the lattice edge), the four Zou-He walls and
corners, and the out-of-range link errors, built with gcc -fsanitize and run
as a child process; any invalid access or undefined operation aborts it.
Runs on the CPU (skipped where gcc or the sanitizer runtimes are missing)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_runs_clean_under_sanitizers(tmp_path):
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("no gcc")
    exe = tmp_path / "oracle_sanitize"
    cmd = [cc, "-std=c99", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           os.path.join(ROOT, "oracle", "d2q9_oracle.c"),
           os.path.join(ROOT, "tests", "oracle_sanitize_main.c"), "-lm", "-o", str(exe)]
    build = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    if build.returncode != 0 and "asan" in (build.stderr + build.stdout).lower():
        pytest.skip("sanitizer runtime not available: " + build.stderr[-300:])
    assert build.returncode == 0, build.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    assert run.returncode == 0, (run.returncode, run.stdout[-2000:], run.stderr[-4000:])
    assert run.stdout.startswith("ok "), run.stdout
