"""Discrete adjoint of the partial-bounce-back channel (lbm_pbb_*).

* FP64 forward on the GPU equals the PyTorch-eager FP64 restatement
  (tests/pbb_reference.py) on a 64x32 downscale: loss, drag history and final
  state within 1e-12 relative.
* FP64 adjoint gradients (dL/ds on every node, dL/du_max) equal torch
  autograd of that restatement within 1e-9 relative to the largest entry.
* Central differences of the FP64 GPU forward on 16 seeded nodes and u_max
  agree with the adjoint within 1e-5 relative (plus 1e-6 of the largest
  sampled gradient, the FP64 round-off floor of the differences) on the
  1024x256 channel at full size.
* The gradient does not depend on the checkpoint interval (bit-identical).
* The FP32 adjoint stays within 1e-4 (relative L2) of the FP64 adjoint.
* A loss whose engine has run another forward refuses its backward; s of
  the wrong shape or on the host is refused before any launch.
"""
import numpy as np
import pytest
import torch

from lbm_amd.adjoint import DragLoss, PbbEngine, cylinder_case
from tests import pbb_reference

pytestmark = pytest.mark.gpu


def small_case(**kw):
    args = dict(nx=64, ny=32, diameter=8.0, cx=14.0, cy=16.0, box=12, noise=0.2, seed=0,
                u_max=0.05, Re=20.0, iterations=120, window=60, every=25, sigma=20.0)
    args.update(kw)
    return cylinder_case(**args)


def _gpu(case, storage="f64", every=None):
    if every is not None:
        case = dict(case, every=every)
    eng = PbbEngine(case, storage=storage)
    dt = torch.float64 if storage == "f64" else torch.float32
    s = torch.tensor(case["s"], dtype=dt, device="cuda")
    loss = eng.forward(s, case["u_max"])
    gs, gu = eng.backward()
    out = dict(loss=loss, drag=eng.drag_history(), final=eng.final_state().double().cpu().numpy(),
               gs=gs.cpu().numpy(), gu=gu)
    eng.close()
    return out


@pytest.fixture(scope="module")
def torch_ref():
    case = small_case()
    s = torch.tensor(case["s"], dtype=torch.float64, requires_grad=True)
    um = torch.tensor(case["u_max"], dtype=torch.float64, requires_grad=True)
    loss, drags, final = pbb_reference.run(case, s, um)
    loss.backward()
    return case, dict(loss=float(loss), drag=drags.detach().numpy(), final=final.detach().numpy(),
                      gs=s.grad.numpy(), gu=float(um.grad))


def test_forward_matches_torch(torch_ref):
    case, ref = torch_ref
    g = _gpu(case)
    assert abs(g["loss"] - ref["loss"]) <= 1e-12 * abs(ref["loss"])
    assert np.abs(g["drag"] - ref["drag"]).max() <= 1e-12 * np.abs(ref["drag"]).max()
    assert np.abs(g["final"] - ref["final"]).max() <= 1e-12


def test_adjoint_matches_torch_autograd(torch_ref):
    case, ref = torch_ref
    g = _gpu(case)
    scale = np.abs(ref["gs"]).max()
    assert scale > 0
    assert np.abs(g["gs"] - ref["gs"]).max() <= 1e-9 * scale
    assert abs(g["gu"] - ref["gu"]) <= 1e-9 * abs(ref["gu"])


def test_checkpoint_interval_does_not_change_gradient():
    case = small_case()
    a = _gpu(case, every=7)
    b = _gpu(case, every=1000)
    assert np.array_equal(a["gs"], b["gs"]) and a["gu"] == b["gu"]


def test_autograd_function():
    case = small_case()
    eng = PbbEngine(case, storage="f64")
    s = torch.tensor(case["s"], dtype=torch.float64, device="cuda", requires_grad=True)
    um = torch.tensor(case["u_max"], dtype=torch.float64, requires_grad=True)
    loss = DragLoss.apply(s, um, eng)
    (2.0 * loss).backward()
    gs, gu = eng.backward()
    assert torch.allclose(s.grad, 2.0 * gs) and abs(float(um.grad) - 2.0 * gu) < 1e-15
    eng.close()


def test_central_differences_full_size():
    """Channel 1024x256, disk D=40 at (200,128) in a noisy 60x60 box, 2000
    iterations, loss over the last 1000, FP64: 16 seeded nodes of the box
    outside the disk and u_max, central differences against the adjoint."""
    case = cylinder_case()
    eng = PbbEngine(case, storage="f64")
    s = torch.tensor(case["s"], dtype=torch.float64, device="cuda")
    eng.forward(s, case["u_max"])
    gs, gu = eng.backward()
    gs = gs.cpu().numpy()
    # nodes of the noisy box outside the disk (inside it s = 1 and the
    # gradient is exactly zero)
    sc = np.asarray(case["s"])
    cand = np.argwhere((sc > 0.0) & (sc < 1.0))
    rng = np.random.default_rng(7)
    nodes = [tuple(int(v) for v in cand[k]) for k in rng.choice(len(cand), 16, replace=False)]
    # central differences: truncation O(eps^2); FP64 round-off of a loss summed
    # over 1000 iterations, divided by 2 eps, leaves ~1e-11 absolute noise, so
    # nodes with near-zero gradient are held to 1e-6 of the largest one
    eps = 1e-5
    rows = []
    for (i, j) in nodes:
        sp = s.clone()
        sp[i, j] += eps
        sm = s.clone()
        sm[i, j] -= eps
        fd = (eng.forward(sp, case["u_max"]) - eng.forward(sm, case["u_max"])) / (2 * eps)
        rows.append((i, j, fd, float(gs[i, j])))
        print(f"node ({i},{j}) s={float(s[i, j]):.4f} fd={fd:.10e} adjoint={gs[i, j]:.10e}",
              flush=True)
    scale = max(abs(r[3]) for r in rows)
    for (i, j, fd, ad) in rows:
        assert abs(fd - ad) <= 1e-5 * abs(ad) + 1e-6 * scale, (i, j, fd, ad)
    e = 1e-7
    fd = (eng.forward(s, case["u_max"] + e) - eng.forward(s, case["u_max"] - e)) / (2 * e)
    print(f"u_max fd={fd:.10e} adjoint={gu:.10e}", flush=True)
    assert abs(fd - gu) <= 1e-5 * abs(fd)
    eng.close()


def test_fp32_adjoint_close_to_fp64():
    case = small_case(nx=128, ny=48, diameter=12.0, cx=30.0, cy=24.0, box=18, iterations=300,
                      window=150, every=40)
    a = _gpu(case, "f64")
    b = _gpu(case, "f32")
    rel = np.linalg.norm(b["gs"] - a["gs"]) / np.linalg.norm(a["gs"])
    print(f"f32 vs f64: loss {abs(b['loss'] - a['loss']) / abs(a['loss']):.3e} "
          f"grad_s {rel:.3e} grad_umax {abs(b['gu'] - a['gu']) / abs(a['gu']):.3e}", flush=True)
    assert abs(b["loss"] - a["loss"]) <= 1e-5 * abs(a["loss"])
    assert rel <= 1e-4, rel
    assert abs(b["gu"] - a["gu"]) <= 1e-4 * abs(a["gu"])


def test_backward_of_a_stale_loss_raises():
    case = small_case(iterations=40, window=20, every=10)
    eng = PbbEngine(case, storage="f64")
    s = torch.tensor(case["s"], dtype=torch.float64, device="cuda", requires_grad=True)
    um = torch.tensor(case["u_max"], dtype=torch.float64, requires_grad=True)
    first = DragLoss.apply(s, um, eng)
    second = DragLoss.apply(s, um, eng)
    with pytest.raises(RuntimeError, match="another forward"):
        first.backward()
    second.backward()      # the engine's last forward: fine
    eng.close()


def test_forward_validates_s():
    case = small_case(iterations=40, window=20, every=10)
    eng = PbbEngine(case, storage="f64")
    with pytest.raises(TypeError):
        eng.forward(torch.tensor(case["s"], dtype=torch.float64), case["u_max"])
    with pytest.raises(ValueError):
        eng.forward(torch.zeros((case["nx"] - 1, case["ny"]), dtype=torch.float64,
                                device="cuda"), case["u_max"])
    eng.close()
