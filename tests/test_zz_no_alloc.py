"""The allocation-free drag read (lbm_drag_read_noalloc) against lbm_drag_read,
both called through the C ABI on the same fused domain: bit-identical sums on
two calls in a row (the second reading fewer iterations), a read that fills
most of the library's fixed staging array, and a clear error for a read that
would not fit in it (after which the entry point still works)."""
import ctypes

import numpy as np
import pytest
import torch

from lbm_amd import _lib
from lbm_amd.domain import FusedDomain, pack_inlets

pytestmark = pytest.mark.gpu

C = np.array([[0, 0], [1, 0], [-1, 0], [0, 1], [0, -1], [1, 1], [-1, -1], [-1, 1], [1, -1]])
MAX = 1048576          # LBM_NOALLOC_DRAG_MAX of include/lbm_d2q9.h


def _domain(nx, ny, block, per_link_obstacles, iterations, ring):
    """Channel with one solid block; its links as one obstacle, or one
    obstacle per link (to make n_obs large).  Runs `iterations` iterations
    and the drag terms of the last one."""
    solid = np.zeros((nx, ny), bool)
    (x0, x1), (y0, y1) = block
    solid[x0:x1, y0:y1] = True
    links = np.array([[i, j, q] for i in range(2, nx - 2) for j in range(2, ny - 2)
                      if not solid[i, j] for q in range(1, 9)
                      if solid[i + C[q, 0], j + C[q, 1]]], dtype=np.int64)
    dist = np.full(len(links), 0.3)
    obs = [(links[k:k + 1], dist[k:k + 1]) for k in range(len(links))] \
        if per_link_obstacles else [(links, dist)]
    bc = [_lib.BC_VELOCITY, _lib.BC_VELOCITY, _lib.BC_VELOCITY, _lib.BC_PRESSURE]
    dom = FusedDomain(nx, ny, 1 / 0.6, 1 / (0.25 / 0.1 + 0.5), bc, obs, ibb=True,
                      storage="f64", ring=ring)
    ul = np.zeros((2, ny))
    ul[0] = 0.04 * 4.0 * np.arange(ny) * (ny - 1 - np.arange(ny)) / (ny - 1) ** 2
    row = (np.zeros((2, nx)), ul, np.zeros((2, nx)), np.zeros((2, ny)), np.ones(ny))
    pool, offs = pack_inlets([row] * ring, nx, ny)
    dom.set_inlets(0, offs, pool)
    dom.init(None)
    dom.steps(1, iterations - 1)
    dom.drag_only(iterations - 1)
    torch.cuda.synchronize()
    return dom


def _read(dom, fn, it0, n):
    out = np.full((n, dom.n_obs, 2), np.nan)
    _lib.check(fn(dom.h, it0, n, out.ctypes.data_as(ctypes.c_void_p), _lib.stream_ptr()),
               fn.__name__)
    return out


def _same(dom, it0, n):
    L = dom.L
    a = _read(dom, L.lbm_drag_read, it0, n)
    b = _read(dom, L.lbm_drag_read_noalloc, it0, n)
    assert np.isfinite(a).all() and np.abs(a).max() > 0.0
    assert np.array_equal(a, b), (it0, n, np.abs(a - b).max())


def test_noalloc_drag_read_equals_drag_read():
    print("case: lbm_drag_read_noalloc, one obstacle, 96x48, 120 iterations", flush=True)
    dom = _domain(96, 48, ((30, 38), (20, 28)), False, 120, 130)
    try:
        _same(dom, 1, 119)         # first call
        _same(dom, 40, 11)         # second call, smaller
    finally:
        dom.close()


def test_noalloc_drag_read_near_capacity_and_rejection():
    print("case: lbm_drag_read_noalloc, one obstacle per link, 160x80, 1260 iterations",
          flush=True)
    dom = _domain(160, 80, ((40, 90), (20, 60)), True, 1260, 1300)
    try:
        n_obs = dom.n_obs
        n_fit = MAX // (2 * n_obs)
        assert 300 < n_fit < 1259 and n_obs > 400
        _same(dom, 1, n_fit)               # fills all but < 2 n_obs doubles of the array
        _same(dom, 500, 25)                # then a smaller read
        with pytest.raises(_lib.LbmError, match="exceeds the staging array"):
            _read(dom, dom.L.lbm_drag_read_noalloc, 1, n_fit + 1)
        _same(dom, 7, 3)                   # still works after the refusal
    finally:
        dom.close()
