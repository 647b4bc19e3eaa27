"""Per-call kernels (one per reference function) against the oracle on
random populations: bit-identical outputs."""
import numpy as np
import pytest
import torch

from lbm_amd import _lib
from oracle import oracle

pytestmark = pytest.mark.gpu

NX, NY = 37, 29


def _rand(seed):
    rng = np.random.default_rng(seed)
    w = np.array([4. / 9.] + [1. / 9.] * 4 + [1. / 36.] * 4)
    g = w[:, None, None] * (1.0 + 0.2 * rng.standard_normal((9, NX, NY)))
    return rng, np.ascontiguousarray(g)


def _d(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _st():
    return _lib.stream_ptr()


def test_macro_equilibrium_col_str():
    L = _lib.load()
    rng, g = _rand(1)
    rho, u = np.zeros((NX, NY)), np.zeros((2, NX, NY))
    oracle.macro(g, rho, u)
    dg, drho, du = _d(g), _d(np.zeros((NX, NY))), _d(np.zeros((2, NX, NY)))
    _lib.check(L.lbm_macro(NX, NY, dg.data_ptr(), drho.data_ptr(), du.data_ptr(), _st()))
    assert np.array_equal(drho.cpu().numpy(), rho)
    assert np.array_equal(du.cpu().numpy(), u)
    geq = np.zeros_like(g)
    oracle.equilibrium(u, rho, geq)
    dgeq = _d(np.zeros_like(g))
    _lib.check(L.lbm_equilibrium(NX, NY, du.data_ptr(), drho.data_ptr(), dgeq.data_ptr(), _st()))
    assert np.array_equal(dgeq.cpu().numpy(), geq)
    gup = np.zeros_like(g)
    g2 = g.copy()
    oracle.col_str(g2, geq, gup, 1.0 / 0.62, 1.0 / 0.9)
    dgup = _d(np.zeros_like(g))
    _lib.check(L.lbm_col_str(NX, NY, dg.data_ptr(), dgeq.data_ptr(), dgup.data_ptr(),
                             1.0 / 0.62, 1.0 / 0.9, _st()))
    assert np.array_equal(dgup.cpu().numpy(), gup)
    assert np.array_equal(dg.cpu().numpy(), g2)


@pytest.mark.parametrize("kind", ["left", "right", "right_pressure", "top", "bottom",
                                  "bottom_left", "top_left", "top_right", "bottom_right"])
def test_zou_he(kind):
    L = _lib.load()
    rng, g = _rand(2)
    rho = 1.0 + 0.01 * rng.standard_normal((NX, NY))
    u = 0.05 * rng.standard_normal((2, NX, NY))
    n = NY if kind in ("left", "right", "right_pressure") else NX
    side = 0.05 * rng.standard_normal((2, n))
    rr = 1.0 + 0.01 * rng.standard_normal(NY)
    go, ro, uo = g.copy(), rho.copy(), u.copy()
    oracle.zou_he(kind, uo, ro, go, side=side, rho_right=rr)
    dg, dr, du = _d(g), _d(rho), _d(u)
    dside, drr = _d(side), _d(rr)   # keep the inputs alive until the kernel has run
    st = _st()
    if kind in ("bottom_left", "top_left", "top_right", "bottom_right"):
        rc = getattr(L, f"lbm_zou_he_{kind}_corner")(NX, NY, du.data_ptr(), dr.data_ptr(),
                                                     dg.data_ptr(), st)
    elif kind == "right_pressure":
        rc = L.lbm_zou_he_right_wall_pressure(NX, NY, du.data_ptr(), drr.data_ptr(),
                                              dside.data_ptr(), dr.data_ptr(), dg.data_ptr(), st)
    else:
        name = {"left": "left", "right": "right", "top": "top", "bottom": "bottom"}[kind]
        rc = getattr(L, f"lbm_zou_he_{name}_wall_velocity")(NX, NY, du.data_ptr(),
                                                            dside.data_ptr(), dr.data_ptr(),
                                                            dg.data_ptr(), st)
    _lib.check(rc)
    assert np.array_equal(dg.cpu().numpy(), go)
    assert np.array_equal(dr.cpu().numpy(), ro)
    assert np.array_equal(du.cpu().numpy(), uo)


@pytest.mark.parametrize("ibb", [0, 1])
def test_bounce_back_and_drag(ibb):
    L = _lib.load()
    rng, g = _rand(3)
    _, gup = _rand(4)
    # links around a 5x5 block at (15..19, 10..14); q points fluid -> solid
    solid = np.zeros((NX, NY), bool)
    solid[15:20, 10:15] = True
    C = np.array([[0, 0], [1, 0], [-1, 0], [0, 1], [0, -1], [1, 1], [-1, -1], [-1, 1], [1, -1]])
    bnd = []
    for i in range(1, NX - 1):
        for j in range(1, NY - 1):
            if solid[i, j]:
                continue
            for q in range(1, 9):
                if solid[i + C[q, 0], j + C[q, 1]]:
                    bnd.append([i, j, q])
    bnd = np.array(bnd, dtype=np.int64)
    p = rng.uniform(0.05, 0.95, len(bnd))
    go = g.copy()
    oracle.bounce_back_obstacle(ibb, bnd, p, gup, go)
    dg, dbnd, dp, dgup = _d(g), _d(bnd), _d(p), _d(gup)
    _lib.check(L.lbm_bounce_back_obstacle(NX, NY, ibb, len(bnd), dbnd.data_ptr(),
                                          dp.data_ptr(), dgup.data_ptr(), dg.data_ptr(), _st()))
    assert np.array_equal(dg.cpu().numpy(), go)
    out = _d(np.zeros(2))
    _lib.check(L.lbm_drag_lift(NX, NY, len(bnd), dbnd.data_ptr(), dgup.data_ptr(),
                               dg.data_ptr(), out.data_ptr(), _st()))
    assert tuple(out.cpu().numpy()) == oracle.drag_sums(bnd, gup, go)
