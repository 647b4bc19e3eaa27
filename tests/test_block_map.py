"""The work-group -> unit map that deals blocks of consecutive tiles (batch
items, adjoint node blocks) to the XCDs, lbm::xcd_block_unit in
lbm_amd/csrc/d2q9_common.h, checked on the host by tests/block_map_main.cpp:
a permutation for every size and block shift, whole blocks per XCD in order,
the tail and bs = 0 unchanged.  CPU only (built with hipcc for the host;
skipped where hipcc is missing)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_xcd_block_unit_is_a_blocked_permutation(tmp_path):
    hipcc = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)
    if hipcc is None:
        pytest.skip("no hipcc")
    exe = tmp_path / "block_map"
    cmd = [hipcc, "-std=c++17", "-O1", os.path.join(ROOT, "tests", "block_map_main.cpp"),
           "-o", str(exe)]
    build = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr[-3000:]
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, (run.returncode, run.stdout[-2000:])
    assert run.stdout.startswith("ok "), run.stdout
