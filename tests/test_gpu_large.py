"""Lattices whose state buffer exceeds 128 MB, where the step sweeps its tiles
last first on odd iterations (smaller lattices, which the other tests use,
keep one direction): FP64 bit for bit against the oracle on every field and
every drag sum, and the 16-bit forms within the tolerances of
tests/test_gpu_forms.py (velocity and density at every node, so that a tile
the sweep missed or ran twice shows), over seven iterations of an
obstacle-array channel (squares of side 20-60, IBB links)."""
import numpy as np
import pytest

from tests.cases import array_domain_case, array_oracle_case
from tests.test_gpu_forms import TOL

pytestmark = pytest.mark.gpu

IT = 7
BYTES = {"f64": 72, "f16": 36, "f16e8": 40}


def _case(form, ny):
    out = array_domain_case(form, IT, ny=ny, n_obs=6, seed=2, side=(20, 60))
    nx = 4 * ny
    assert nx * ny * BYTES[form] > 128 << 20, "lattice too small for the alternating sweep"
    return out


def test_fp64_alternating_sweep_matches_oracle():
    print("case f64 2816x704 (state 143 MB)", flush=True)
    g = _case("f64", 704)
    o = array_oracle_case(g, IT)
    for k in ("rho", "u", "g", "g_up", "drag"):
        assert np.array_equal(g[k], o[k]), k


@pytest.mark.parametrize("form", ["f16", "f16e8"])
def test_16bit_alternating_sweep_close_to_oracle(form):
    print(f"case {form} 4096x1024", flush=True)
    g = _case(form, 1024)
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
