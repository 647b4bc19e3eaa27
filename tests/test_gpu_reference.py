"""The reference's own tests and the configurations' stated sizes, run on the
HIP path through the drop-in ``lbm_amd.lattice`` + ``lbm_amd.run``.

* lbm/tst/cavity/test_cavity.py: default cavity (100x100, 10 001
  iterations); the four asserted values to 1e-6 and all 200 values of the
  golden profiles cavity_{vx,uy}_ref to 5e-6, and the whole final state bit
  for bit against the oracle.
* lbm/tst/poiseuille/test_poiseuille.py: L1 error < 1e-3, and the final
  state bit for bit against the oracle.
* Cavity 200x200, 20 001 iterations (it_max = 20 000), FP64, bit for bit.
* Schaefer-Turek channel at full size (1073x200, IBB cylinder), 2000
  iterations: Cd/Cl of every iteration and the final state bit for bit.
"""
import os

import numpy as np
import pytest

from tests.cases import (cavity200_app, ref_cavity, ref_poiseuille, run_gpu, run_oracle,
                         turek200_app)

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FIELDS = ("rho", "u", "g", "g_up", "g_eq")


def _same(a, b, what):
    for k in FIELDS:
        assert a[k].shape == b[k].shape, (what, k)
        if not np.array_equal(a[k], b[k]):
            d = np.abs(a[k] - b[k])
            idx = np.unravel_index(np.argmax(d), d.shape)
            raise AssertionError(f"{what}: {k} differs, max {d.max():.3e} at {idx}")
    if a["drag"] is not None:
        assert np.array_equal(a["drag"], b["drag"]), (what, "drag")


def test_reference_cavity_on_hip_path():
    print("case: reference cavity 100x100, 10001 iterations", flush=True)
    gpu, app, lat = run_gpu(ref_cavity, 10001, chunk=1000)
    vx, uy = app.line_fields(lat)
    assert abs(vx[10] - 0.12493089684236539) < 1.0e-6
    assert abs(vx[50] - 0.05295104908939561) < 1.0e-6
    assert abs(uy[10] + 0.05968571489630510) < 1.0e-6
    assert abs(uy[50] + 0.19792323493599165) < 1.0e-6
    ref_vx = np.loadtxt(os.path.join(GOLD, "cavity_vx_ref"))
    ref_uy = np.loadtxt(os.path.join(GOLD, "cavity_uy_ref"))
    assert np.abs(vx - ref_vx[:, 1]).max() < 5e-6
    assert np.abs(uy - ref_uy[:, 1]).max() < 5e-6
    ora = run_oracle(ref_cavity, 10001)[0]
    _same(gpu, ora, "reference cavity")


def test_reference_poiseuille_on_hip_path():
    print("case: reference poiseuille 150x50, 11251 iterations", flush=True)
    gpu, app, lat = run_gpu(ref_poiseuille, 11251, chunk=1000)
    assert app.compute_error(lat) < 1.0e-3
    ora = run_oracle(ref_poiseuille, 11251)[0]
    _same(gpu, ora, "reference poiseuille")


def test_cavity200_full_run_bitwise():
    print("case: cavity 200x200, 20001 iterations, FP64", flush=True)
    gpu = run_gpu(cavity200_app, 20001, chunk=1000)[0]
    ora = run_oracle(cavity200_app, 20001)[0]
    _same(gpu, ora, "cavity200")


def test_turek200_full_size_bitwise():
    print("case: turek 1073x200, 2000 iterations, FP64", flush=True)
    gpu = run_gpu(turek200_app, 2000, chunk=1000)[0]
    ora = run_oracle(turek200_app, 2000)[0]
    assert gpu["drag"].shape == (2000, 1, 2)
    # the flow has reached the cylinder: the drag is not round-off
    assert abs(gpu["drag"][-1, 0, 0]) > 1e-4
    _same(gpu, ora, "turek200")
