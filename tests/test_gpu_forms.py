"""Float storage forms of the fused step against the FP64 oracle.

F32 stores f - w in float, F16/BF16 store f - w in half/bfloat16, F16E stores
f - w - E(v) in half with the node's velocity v in a half pair, Q16 stores
f - w in 16-bit fixed point over a range set by 1.5 u_lid, F16E8 is F16E
with the velocity as an 8-bit fixed-point pair; all compute in float.
Tolerances, about three times the largest error measured on these two cases (velocity scale u_ref = max |u| of the oracle, drag scale = max
|Cd| of the oracle over the run):
  F32:   |u - u64| <= 3e-6 u_ref,    |rho - rho64| <= 1e-6,    |Cd - Cd64| <= 3e-6 Cd_ref
  F16:   |u - u64| <= 3e-3 u_ref,    |rho - rho64| <= 2e-4,    |Cd - Cd64| <= 2e-3 Cd_ref
  BF16:  |u - u64| <= 2e-2 u_ref,    |rho - rho64| <= 2e-3,    |Cd - Cd64| <= 8e-3 Cd_ref
  F16E:  |u - u64| <= 1.5e-3 u_ref,  |rho - rho64| <= 1.5e-4,  |Cd - Cd64| <= 1e-3 Cd_ref
  Q16:   |u - u64| <= 8e-4 u_ref,    |rho - rho64| <= 8e-5,    |Cd - Cd64| <= 8e-4 Cd_ref
  F16E8: |u - u64| <= 1.2e-3 u_ref,  |rho - rho64| <= 1.5e-4,  |Cd - Cd64| <= 1e-3 Cd_ref
(60 iterations of an impulsively started flow, the hardest regime for the
reduced forms; long-run statistics are compared by bench.py)."""
import numpy as np
import pytest

from tests.cases import run_gpu, run_oracle, small_cavity, small_channel

pytestmark = pytest.mark.gpu

TOL = {"f32": (3e-6, 1e-6, 3e-6), "f16": (3e-3, 2e-4, 2e-3), "bf16": (2e-2, 2e-3, 8e-3),
       "f16e": (1.5e-3, 1.5e-4, 1e-3), "q16": (8e-4, 8e-5, 8e-4), "f16e8": (1.2e-3, 1.5e-4, 1e-3)}


@pytest.mark.parametrize("form", ["f32", "f16", "bf16", "f16e", "q16", "f16e8"])
@pytest.mark.parametrize("make", [small_cavity, small_channel])
def test_form_close_to_fp64(make, form):
    print(f"case {make.__name__} {form}", flush=True)
    o = run_oracle(make, 60)[0]
    g = run_gpu(make, 60, precision=form)[0]
    tu, tr, td = TOL[form]
    uref = np.abs(o["u"]).max()
    print(f"  max u error / u_ref {np.abs(g['u'] - o['u']).max() / uref:.3e}, "
          f"rho error {np.abs(g['rho'] - o['rho']).max():.3e}", flush=True)
    assert np.abs(g["u"] - o["u"]).max() <= tu * uref
    assert np.abs(g["rho"] - o["rho"]).max() <= tr
    if o["drag"] is not None:
        dref = np.abs(o["drag"][..., 0]).max()
        print(f"  max drag error / Cd_ref {np.abs(g['drag'] - o['drag']).max() / dref:.3e}",
              flush=True)
        assert np.abs(g["drag"] - o["drag"]).max() <= td * dref
