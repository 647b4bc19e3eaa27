"""Lattices whose state buffer holds 128-512 MB, where the step sweeps its
tiles last first on odd iterations (smaller lattices, which the other tests
use, keep one direction and deal tiles in blocks of 4; the 16-bit cases here
also deal whole columns to the XCDs): FP64 bit for bit against the oracle on
every field and every drag sum, and the 16-bit forms within the tolerances of
tests/test_gpu_forms.py (velocity and density at every node, so that a tile
the sweep missed or ran twice shows), over seven iterations of an
obstacle-array channel (squares of side 20-60, IBB links)."""
import numpy as np
import pytest

from tests.cases import array_domain_case, array_oracle_case
from tests.test_gpu_forms import TOL

pytestmark = pytest.mark.gpu

IT = 7
STATE_BYTES = {"f64": 72, "f16": 18, "f16e8": 20}   # per node, one state buffer


def _case(form, ny, aspect=4):
    out = array_domain_case(form, IT, ny=ny, n_obs=6, seed=2, side=(20, 60), aspect=aspect)
    state = aspect * ny * ny * STATE_BYTES[form]
    assert (128 << 20) < state <= (512 << 20), "lattice outside the alternating-sweep range"
    return out


def test_fp64_alternating_sweep_matches_oracle():
    print("case f64 2816x704 (state 143 MB)", flush=True)
    g = _case("f64", 704)
    o = array_oracle_case(g, IT)
    for k in ("rho", "u", "g", "g_up", "drag"):
        assert np.array_equal(g[k], o[k]), k


@pytest.mark.parametrize("form", ["f16", "f16e8"])
def test_16bit_alternating_sweep_close_to_oracle(form):
    print(f"case {form} 8192x1024", flush=True)
    g = _case(form, 1024, aspect=8)
    o = array_oracle_case(g, IT)
    tu, tr, td = TOL[form]
    uref = np.abs(o["u"]).max()
    print(f"  max u error / u_ref {np.abs(g['u'] - o['u']).max() / uref:.3e}, "
          f"rho error {np.abs(g['rho'] - o['rho']).max():.3e}", flush=True)
    assert np.abs(g["u"] - o["u"]).max() <= tu * uref
    assert np.abs(g["rho"] - o["rho"]).max() <= tr
    # after seven iterations the inflow has not reached the obstacles: their drag
    # sums are round-off (~1e-14), hence the absolute floor
    dref = np.abs(o["drag"][..., 0]).max()
    assert np.abs(g["drag"] - o["drag"]).max() <= td * dref + 1e-12
