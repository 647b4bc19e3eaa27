"""Float storage forms of the fused step against the FP64 oracle.

F32 stores f - w in float, F16/BF16 store f - w in half/bfloat16, F16E stores
f - w - E(v) in half with the node's velocity v in a half pair; all four
compute in float.  Tolerances (velocity scale u_ref = max |u| of the oracle,
drag scale = max |Cd| of the oracle over the run):
  F32:  |u - u64| <= 2e-5 u_ref,  |rho - rho64| <= 2e-6,  |Cd - Cd64| <= 1e-4 Cd_ref
  F16:  |u - u64| <= 2e-2 u_ref,  |rho - rho64| <= 2e-3,  |Cd - Cd64| <= 2e-2 Cd_ref
  BF16: |u - u64| <= 1e-1 u_ref,  |rho - rho64| <= 1e-2,  |Cd - Cd64| <= 1e-1 Cd_ref
  F16E: |u - u64| <= 2e-3 u_ref,  |rho - rho64| <= 2e-4,  |Cd - Cd64| <= 2e-3 Cd_ref
  Q16:  |u - u64| <= 2e-3 u_ref,  |rho - rho64| <= 5e-4,  |Cd - Cd64| <= 2e-3 Cd_ref
        (uniform resolution over the range of f - w set by 1.5 u_lid: ~2e-6
        per population here, where the other forms resolve small values finer)
(60 iterations of an impulsively started flow, the hardest regime for the
reduced forms; long-run statistics are compared by bench.py)."""
import numpy as np
import pytest

from tests.cases import run_gpu, run_oracle, small_cavity, small_channel

pytestmark = pytest.mark.gpu

TOL = {"f32": (2e-5, 2e-6, 1e-4), "f16": (2e-2, 2e-3, 2e-2), "bf16": (1e-1, 1e-2, 1e-1),
       "f16e": (2e-3, 2e-4, 2e-3), "q16": (2e-3, 5e-4, 2e-3)}


@pytest.mark.parametrize("form", ["f32", "f16", "bf16", "f16e", "q16"])
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
        assert np.abs(g["drag"] - o["drag"]).max() <= td * dref
