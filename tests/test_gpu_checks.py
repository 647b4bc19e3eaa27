"""Density sums, failure detection and argument checks of the fused step.

* lbm_mass is the FP64 sum of the state's populations: for every storage
  form it equals the double sum of the epilogue's g_up of the same iteration
  (the float forms' populations are w_q + h_q with the exact double weights)
  within 1e-12 relative.
* run() raises InstabilityError when the populations stop being finite (a
  NaN put into the initial g on the host) or drift beyond max_mass_drift.
* lbm_fused_set_inlets refuses offsets whose side array would leave the pool.
"""
import numpy as np
import pytest
import torch

from lbm_amd import _lib
from lbm_amd.domain import FusedDomain, pack_inlets
from tests.cases import run_gpu, small_cavity, small_channel

pytestmark = pytest.mark.gpu

W = np.array([4. / 9.] + [1. / 9.] * 4 + [1. / 36.] * 4)


def _channel_domain(storage, nx=128, ny=64, iterations=30, seed=3):
    rng = np.random.default_rng(seed)
    g0 = W[:, None, None] * (1.0 + 0.05 * rng.standard_normal((9, nx, ny)))
    solid = np.zeros((nx, ny), bool)
    solid[40:48, 28:36] = True
    C = np.array([[0, 0], [1, 0], [-1, 0], [0, 1], [0, -1], [1, 1], [-1, -1], [-1, 1], [1, -1]])
    links = np.array([[i, j, q] for i in range(1, nx - 1) for j in range(1, ny - 1)
                      if not solid[i, j] for q in range(1, 9)
                      if solid[i + C[q, 0], j + C[q, 1]]], dtype=np.int64)
    bc = [_lib.BC_VELOCITY, _lib.BC_VELOCITY, _lib.BC_VELOCITY, _lib.BC_PRESSURE]
    dom = FusedDomain(nx, ny, 1 / 0.6, 1 / (0.25 / 0.1 + 0.5), bc,
                      [(links, np.full(len(links), 0.3))], ibb=True, storage=storage,
                      ring=iterations + 2)
    ul = np.zeros((2, ny))
    ul[0] = 0.03
    rows = [(np.zeros((2, nx)), ul, np.zeros((2, nx)), np.zeros((2, ny)), np.ones(ny))]
    dom.set_inlet_rows(0, rows * (iterations + 1))
    dom.init(torch.from_numpy(np.ascontiguousarray(g0)).cuda())
    for it in range(1, iterations):
        dom.step(it)
    return dom, iterations - 1


@pytest.mark.parametrize("storage", ["f64", "f32", "f16", "bf16", "f16e"])
def test_mass_is_the_fp64_sum_of_the_state(storage):
    print(f"case: lbm_mass {storage}", flush=True)
    dom, last = _channel_domain(storage)
    m = dom.mass(last)
    g = torch.empty((9, dom.nx, dom.ny), dtype=torch.float64, device="cuda")
    gup = torch.empty_like(g)
    dom.epilogue(last, g, gup)
    ref = float(np.sum(gup.cpu().numpy(), dtype=np.float64))
    dom.close()
    assert abs(m - ref) <= 1e-12 * abs(ref), (m, ref)


def test_mass_of_the_rest_state_has_no_weight_bias():
    """A float form at rest stores h = 0: its density sum is N sum_q w_q in
    double, with no float-rounded weights (which would add 7.45e-9 per node)."""
    nx, ny = 256, 128
    bc = [_lib.BC_VELOCITY] * 4
    dom = FusedDomain(nx, ny, 1 / 0.7, 1 / (0.25 / 0.2 + 0.5), bc, storage="f16", ring=4)
    z = (np.zeros((2, nx)), np.zeros((2, ny)), np.zeros((2, nx)), np.zeros((2, ny)),
         np.zeros(ny))
    dom.set_inlet_rows(0, [z] * 3)
    dom.init(None)
    dom.step(1)
    m = dom.mass(1)
    dom.close()
    assert abs(m - nx * ny * float(np.sum(W))) <= 1e-9, m


class _NaNStart:
    def initialize(self, lattice):
        super().initialize(lattice)
        g = lattice.g
        g[1, 7, 9] = np.nan
        lattice.g = g


def test_run_raises_on_non_finite_populations():
    from lbm_amd.run import InstabilityError

    def make():
        app = small_cavity()
        app.__class__ = type("nan_cavity", (_NaNStart, type(app)), {})
        return app
    with pytest.raises(InstabilityError, match="non-finite"):
        run_gpu(make, 60, chunk=16)


def test_run_records_finite_mass_and_checks_drift():
    from lbm_amd.run import InstabilityError
    out, app, lat = run_gpu(small_channel, 60, chunk=16)
    hist = np.array([m for _, m in lat.mass_history])
    assert len(hist) >= 3 and np.all(np.isfinite(hist))

    def make():
        app = small_channel()
        app.max_mass_drift = 1e-30       # any change at all
        return app
    with pytest.raises(InstabilityError, match="drift"):
        run_gpu(make, 60, chunk=16)


def test_set_inlets_rejects_offsets_past_the_pool():
    nx, ny = 64, 32
    bc = [_lib.BC_VELOCITY, _lib.BC_VELOCITY, _lib.BC_VELOCITY, _lib.BC_PRESSURE]
    dom = FusedDomain(nx, ny, 1 / 0.6, 1 / (0.25 / 0.1 + 0.5), bc, storage="f64", ring=8)
    row = (np.zeros((2, nx)), np.zeros((2, ny)), np.zeros((2, nx)), np.zeros((2, ny)),
           np.ones(ny))
    pool, offs = pack_inlets([row])
    dom.set_inlets(0, offs, pool)                   # valid
    bad = offs.copy()
    bad[0, 1] = pool.size - ny                      # u_left needs 2*ny doubles
    with pytest.raises(_lib.LbmError, match="exceeds the pool"):
        dom.set_inlets(0, bad, pool)
    bad = offs.copy()
    bad[0, 4] = -1
    with pytest.raises(_lib.LbmError):
        dom.set_inlets(0, bad, pool)
    with pytest.raises(ValueError, match="u_left"):
        dom.set_inlet_rows(0, [(row[0], np.zeros((2, ny - 1)), row[2], row[3], row[4])])
    dom.close()


def test_amplitude_tables_check_their_arguments():
    """lbm_fused_set_amplitudes needs lbm_fused_set_profiles first; a batch
    case's wall inputs belong to the batch."""
    from lbm_amd.batch import Batch, cavity_ensemble
    nx, ny = 64, 32
    dom = FusedDomain(nx, ny, 1 / 0.6, 1 / (0.25 / 0.1 + 0.5), [_lib.BC_VELOCITY] * 4,
                      storage="f64", ring=8)
    with pytest.raises(_lib.LbmError, match="set_profiles first"):
        dom.set_amplitudes(0, np.zeros((3, 5)))
    top = np.zeros((2, nx))
    top[0] = 1.0
    dom.set_profiles([None, None, top, None, None])
    dom.set_amplitudes(0, np.zeros((8, 5)))                   # a whole ring
    with pytest.raises(_lib.LbmError):
        dom.set_amplitudes(0, np.zeros((9, 5)))               # more rows than the ring
    dom.close()
    b = Batch(cavity_ensemble(n_cases=2, seed=1, n_lo=32, n_hi=40), storage="f64", ring=8)
    with pytest.raises(_lib.LbmError, match="batch"):
        b.case(0).set_profiles([None] * 5)
    b.close()


def test_run_stops_when_separable_inlets_disagree_with_set_inlets():
    """An app whose declared amplitude x profile is not what its set_inlets
    writes is caught at the next chunk start, not run with wrong walls."""
    class _Lying:
        def separable_inlets(self, lattice):
            profiles, amp = super().separable_inlets(lattice)
            return profiles, (lambda it: tuple(1.001 * a for a in amp(it)))

        def set_inlets(self, lattice, it):
            super().set_inlets(lattice, it)

    def make():
        app = small_cavity()
        app.__class__ = type("lying_cavity", (_Lying, type(app)), {})
        app.sigma = 1          # the lid is moving from iteration 1
        return app
    with pytest.raises(RuntimeError, match="disagrees"):
        run_gpu(make, 60, chunk=16)
