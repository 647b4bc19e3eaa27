"""Batched ensembles (lbm_batch_*) against the oracle: every case of a batch
advanced by the batched launch equals the same case run alone on the oracle
(the reference loop, run.py:24-54) -- bit for bit in FP64 on rho, u, g, g_up
and the drag sums -- and the FP32 batch stays within 2e-5 of the oracle's
velocity scale.  Cases differ in size (odd sizes included), relaxation rates
and walls; one carries IBB obstacle links.  The 256-cavity ensemble itself
is run whole and six of its cases (smallest, largest, odd sizes, extreme
relaxation rates) are compared with the oracle.  The 16-bit forms, which
have no oracle of their own, are checked for equality between the batched
and the single-domain launch."""
import math

import numpy as np
import pytest
import torch

from lbm_amd import _lib
from lbm_amd.batch import Batch, autotune, cavity_amplitudes, cavity_ensemble
from lbm_amd.domain import FusedDomain, pack_inlets, q16_range
from tests.cases import batch_case_oracle

pytestmark = pytest.mark.gpu

IT = 60


def _cases():
    cases = cavity_ensemble(n_cases=5, seed=3, n_lo=17, n_hi=50)
    for c in cases:
        c["sigma"] = 15
    # a channel with a 5x5 block (plain links listed by hand, IBB p = 0.3)
    nx, ny = 64, 24
    solid = np.zeros((nx, ny), bool)
    solid[20:25, 9:14] = True
    C = np.array([[0, 0], [1, 0], [-1, 0], [0, 1], [0, -1], [1, 1], [-1, -1], [-1, 1], [1, -1]])
    links = np.array([[i, j, q] for i in range(1, nx - 1) for j in range(1, ny - 1)
                      if not solid[i, j] for q in range(1, 9)
                      if solid[i + C[q, 0], j + C[q, 1]]], dtype=np.int64)
    left = np.zeros((2, ny))
    yy = np.arange(ny)
    left[0] = 4.0 * yy * (ny - 1 - yy) / (ny - 1) ** 2
    tau = 0.6
    cases.append(dict(nx=nx, ny=ny, om_p=1 / tau, om_m=1 / (0.25 / (tau - 0.5) + 0.5),
                      bc=[_lib.BC_VELOCITY, _lib.BC_VELOCITY, _lib.BC_VELOCITY, _lib.BC_PRESSURE],
                      profiles=[None, left, None, None, np.ones(ny)], links=links,
                      ibb=np.full(len(links), 0.3), ibb_on=np.ones(len(links), np.uint8),
                      obs_offsets=[0, len(links)], sigma=15, u_lid=0.05, channel=True))
    return cases


def _amp(cases, n):
    amp = np.zeros((n, len(cases), 5))
    for k in range(n):
        for c, cs in enumerate(cases):
            ret = (1.0 - math.exp(-k ** 2 / (2.0 * cs["sigma"] ** 2)))
            if cs.get("channel"):
                amp[k, c, 1] = ret * cs["u_lid"]
                amp[k, c, 4] = 1.0
            else:
                amp[k, c, 2] = cs["u_lid"] * ret
    return amp


def _final(dom, nx, ny, last):
    """Last iteration with its moments, the drag sums of iterations 0..last,
    and the final fields (epilogue)."""
    rho = torch.empty((nx, ny), dtype=torch.float64, device="cuda")
    u = torch.empty((2, nx, ny), dtype=torch.float64, device="cuda")
    g = torch.empty((9, nx, ny), dtype=torch.float64, device="cuda")
    gup = torch.empty_like(g)
    dom.step(last, rho, u)
    drag = None
    if dom.n_obs:
        dom.drag_only(last)
        drag = dom.drag_read(0, last + 1)
    dom.epilogue(last, g, gup, rho, u)
    return dict(rho=rho.cpu().numpy(), u=u.cpu().numpy(), g=g.cpu().numpy(),
                g_up=gup.cpu().numpy(), drag=drag)


def _single(cs, amp_c, storage, vector_width=0, cpt=0):
    """One case alone through FusedDomain with a pool inlet table of the same values."""
    nx, ny = cs["nx"], cs["ny"]
    links = cs.get("links")
    obs = [(links, cs["ibb"])] if links is not None else []
    dom = FusedDomain(nx, ny, cs["om_p"], cs["om_m"], cs["bc"], obs, ibb=links is not None,
                      storage=storage, ring=IT + 2, vector_width=vector_width,
                      cells_per_thread=cpt, q16_range=_q16(storage))
    rows = []
    profs = cs["profiles"]
    for k in range(IT + 1):
        snap = []
        for sd, n in enumerate((2 * nx, 2 * ny, 2 * nx, 2 * ny, ny)):
            shape = (ny,) if sd == 4 else (2, n // 2)
            a = np.zeros(shape) if profs[sd] is None else amp_c[k, sd] * np.asarray(profs[sd])
            snap.append(a)
        rows.append(tuple(snap))
    pool, offs = pack_inlets(rows)
    dom.set_inlets(0, offs, pool)
    dom.init(None)
    for it in range(1, IT - 1):
        dom.step(it)
    out = _final(dom, nx, ny, IT - 1)
    dom.close()
    return out


def _q16(storage):
    """Q16's range of f - w: velocities up to 0.15 (lids of 0.1, channel 0.05)."""
    return q16_range(0.15) if storage == "q16" else None


def _batch(cases, amp, storage, **kw):
    b = Batch(cases, storage=storage, ring=IT + 2, q16_range=_q16(storage), **kw)
    b.set_amplitudes(0, amp[:IT + 1])
    b.init()
    b.steps(1, IT - 2)
    outs = [_final(b.case(c), cs["nx"], cs["ny"], IT - 1) for c, cs in enumerate(cases)]
    b.close()
    return outs


def _same(a, b, what):
    for k in ("rho", "u", "g", "g_up"):
        if not np.array_equal(a[k], b[k]):
            d = np.abs(a[k] - b[k])
            raise AssertionError(f"{what}: {k} differs, max {d.max():.3e}")
    if b.get("drag") is not None:
        assert a["drag"] is not None and np.array_equal(a["drag"], b["drag"]), (what, "drag")


@pytest.mark.parametrize("kw", [dict(), dict(grid=7), dict(cells_per_thread=1, grid=3)])
def test_batch_fp64_matches_oracle(kw):
    print(f"case: batch f64 {kw}", flush=True)
    cases = _cases()
    amp = _amp(cases, IT + 1)
    outs = _batch(cases, amp, "f64", **kw)
    for c, cs in enumerate(cases):
        ref = batch_case_oracle(cs, amp[:, c], IT)
        _same(outs[c], ref, f"case {c} ({cs['nx']}x{cs['ny']})")


def test_batch_fp32_close_to_oracle():
    print("case: batch f32", flush=True)
    cases = _cases()
    amp = _amp(cases, IT + 1)
    outs = _batch(cases, amp, "f32")
    for c, cs in enumerate(cases):
        ref = batch_case_oracle(cs, amp[:, c], IT)
        uref = np.abs(ref["u"]).max()
        assert np.abs(outs[c]["u"] - ref["u"]).max() <= 2e-5 * uref, c
        assert np.abs(outs[c]["rho"] - ref["rho"]).max() <= 2e-6, c


C4_IT = 400
C4_PICKS = (23, 53, 180, 174, 3, 6)   # n = 65, 511, 499 (Re 993), 479 (tau 1.81), 185, 97


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_c4_ensemble_matches_oracle(storage):
    """The 256-cavity ensemble (n ~ U{64..512}, Re ~ logU[100, 1000], seed 0)
    in one batch, 400 iterations of the lid ramp; six cases against the
    oracle: FP64 bit for bit, FP32 within 2e-5 of the velocity scale plus
    the FP32 rounding floor (1e-7)."""
    from lbm_amd.batch import cavity_amplitudes
    print(f"case: 256-cavity ensemble {storage}, {C4_IT} iterations", flush=True)
    cases = cavity_ensemble(n_cases=256, seed=0)
    amp = cavity_amplitudes(cases, 0, C4_IT + 1)
    b = Batch(cases, storage=storage, ring=C4_IT + 2)
    b.set_amplitudes(0, amp)
    b.init()
    b.steps(1, C4_IT - 2)
    outs = {c: _final(b.case(c), cases[c]["nx"], cases[c]["ny"], C4_IT - 1) for c in C4_PICKS}
    b.close()
    for c in C4_PICKS:
        ref = batch_case_oracle(cases[c], amp[:, c], C4_IT)
        if storage == "f64":
            _same(outs[c], ref, f"ensemble case {c}")
        else:
            # 2e-5 of the velocity scale, plus the FP32 rounding floor of a
            # density near 1 (one float ulp of 1.0 is 1.19e-7): at 400
            # iterations the lid of the largest cavities has reached only
            # 3e-4 (ramp sigma = 10 n), where that floor dominates
            uref = np.abs(ref["u"]).max()
            err = np.abs(outs[c]["u"] - ref["u"]).max()
            print(f"case {c}: max |u| {uref:.3e}, max u error {err:.3e}", flush=True)
            assert err <= 2e-5 * uref + 1e-7, c
            assert np.abs(outs[c]["rho"] - ref["rho"]).max() <= 2e-6, c


@pytest.mark.parametrize("storage,vw", [("f16", 1), ("f16", 4), ("bf16", 1), ("bf16", 4),
                                        ("f16e", 1), ("f16e8", 1), ("q16", 1)])
def test_batch_16bit_equals_single_runs(storage, vw):
    """16-bit storage, one or four nodes per work-item: the four-node kernel
    needs ny % 4 == 0 and nx*ny a multiple of 1024 (one group of four per
    work-item), which these cases meet."""
    cases = cavity_ensemble(n_cases=3, seed=5, n_lo=32, n_hi=32)
    cases[1] = cavity_ensemble(n_cases=1, seed=6, n_lo=64, n_hi=64)[0]
    for c in cases:
        c["sigma"] = 15
    ch = _cases()[-1]
    ny = 32
    ch.update(ny=ny, nx=64, links=ch["links"][ch["links"][:, 1] < ny - 1])
    ch["ibb"] = np.full(len(ch["links"]), 0.3)
    ch["ibb_on"] = np.ones(len(ch["links"]), np.uint8)
    ch["obs_offsets"] = [0, len(ch["links"])]
    yy = np.arange(ny)
    left = np.zeros((2, ny))
    left[0] = 4.0 * yy * (ny - 1 - yy) / (ny - 1) ** 2
    ch["profiles"] = [None, left, None, None, np.ones(ny)]
    cases.append(ch)
    amp = _amp(cases, IT + 1)
    b = Batch(cases, storage=storage, ring=IT + 2, vector_width=vw, cells_per_thread=1,
              q16_range=_q16(storage))
    assert b.vector_width == vw
    b.close()
    outs = _batch(cases, amp, storage, vector_width=vw, cells_per_thread=1)
    for c, cs in enumerate(cases):
        ref = _single(cs, amp[:, c], storage, vector_width=vw, cpt=1)
        for k in ("rho", "u", "g", "g_up"):
            assert np.array_equal(outs[c][k], ref[k]), (c, k)


def test_autotune_returns_a_valid_shape(tmp_path):
    cases = _cases()
    best, res = autotune(cases, "f32", steps=4, cache=str(tmp_path / "t.json"))
    assert len(res) >= 3 and best["cells_per_thread"] in (1, 2, 4, 8)
    b = Batch(cases, storage="f32", ring=8, **best)
    b.close()
