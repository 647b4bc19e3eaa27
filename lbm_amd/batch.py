"""Batched ensembles: many lattices advanced by one launch per iteration
(lbm_batch_* of include/lbm_d2q9.h), with an autotuner for the launch shape.

A case is described by a dict: nx, ny, om_p, om_m, bc (four LBM_BC_* codes),
profiles (five arrays or None: u_bot [2][nx], u_left [2][ny], u_top [2][nx],
u_right [2][ny], rho_right [ny]) and optionally links/ibb/ibb_on/obs_offsets.
Wall values are amplitude(iteration, case, side) x profile.
"""
import ctypes
import json
import math
import os
import time

import numpy as np

from lbm_amd import _lib
from lbm_amd.domain import FusedDomain


class Batch:
    def __init__(self, cases, storage="f32", ring=1002, cells_per_thread=0, vector_width=0,
                 grid=0, q16_range=None):
        self.L = _lib.load()
        self.cases = cases
        self.storage = storage
        self.ring = int(ring)
        self._keep = []
        descs = (_lib.CaseDesc * len(cases))()
        for d, c in zip(descs, cases):
            d.nx, d.ny = int(c["nx"]), int(c["ny"])
            d.om_p, d.om_m = float(c["om_p"]), float(c["om_m"])
            for i in range(4):
                d.bc[i] = int(c["bc"][i])
            for sd, prof in enumerate(c.get("profiles", [None] * 5)):
                if prof is not None:
                    a = np.ascontiguousarray(prof, dtype=np.float64)
                    self._keep.append(a)
                    d.profile_host[sd] = a.ctypes.data_as(_lib.c_dp)
            links = c.get("links")
            if links is not None and len(links):
                b = np.ascontiguousarray(links, dtype=np.int64)
                ib = np.ascontiguousarray(c.get("ibb", np.zeros(len(b))), dtype=np.float64)
                md = np.ascontiguousarray(c.get("ibb_on", np.zeros(len(b))), dtype=np.uint8)
                off = np.ascontiguousarray(c.get("obs_offsets", [0, len(b)]), dtype=np.int32)
                self._keep += [b, ib, md, off]
                d.n_links = len(b)
                d.links_host = b.ctypes.data_as(ctypes.POINTER(_lib.c_i64))
                d.ibb_host = ib.ctypes.data_as(_lib.c_dp)
                d.link_ibb_host = md.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
                d.n_obs = len(off) - 1
                d.obs_offsets_host = off.ctypes.data_as(ctypes.POINTER(_lib.c_i32))
        bd = _lib.BatchDesc()
        bd.n_cases = len(cases)
        bd.cases = descs
        bd.storage = _lib.STORE[storage]
        bd.cells_per_thread = int(cells_per_thread)
        bd.vector_width = int(vector_width)
        bd.ring = self.ring
        bd.grid = int(grid)
        if q16_range is not None:
            q = np.ascontiguousarray(q16_range, dtype=np.float64).reshape(9, 2)
            self._keep.append(q)
            bd.q16_range_host = q.ctypes.data_as(_lib.c_dp)
        h = _lib.c_vp()
        _lib.check(self.L.lbm_batch_create(ctypes.byref(bd), ctypes.byref(h)), "lbm_batch_create")
        self.h = h
        info = np.zeros(4, dtype=np.int32)
        _lib.check(self.L.lbm_batch_info(self.h, info.ctypes.data_as(_lib.c_vp)))
        self.n_work, self.grid, self.cpt, self.vector_width = (int(v) for v in info)
        self.cells = int(sum(int(c["nx"]) * int(c["ny"]) for c in cases))

    def close(self):
        if getattr(self, "h", None):
            self.L.lbm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_amplitudes(self, row0, amp):
        a = np.ascontiguousarray(amp, dtype=np.float64)
        assert a.ndim == 3 and a.shape[1:] == (len(self.cases), 5)
        _lib.check(self.L.lbm_batch_set_amplitudes(self.h, int(row0), a.shape[0],
                                                   a.ctypes.data_as(_lib.c_vp), _lib.stream_ptr()),
                   "lbm_batch_set_amplitudes")

    def init(self):
        _lib.check(self.L.lbm_batch_init(self.h, _lib.stream_ptr()), "lbm_batch_init")

    def steps(self, it0, n):
        _lib.check(self.L.lbm_batch_steps(self.h, int(it0), int(n), _lib.stream_ptr()),
                   "lbm_batch_steps")

    def prepare(self, it0, n):
        """Capture and upload the graph steps(it0, n) will replay, without running it."""
        _lib.check(self.L.lbm_batch_steps_prepare(self.h, int(it0), int(n), _lib.stream_ptr()),
                   "lbm_batch_steps_prepare")

    def case(self, c):
        """A non-owning FusedDomain view of case c (step, epilogue, mass, drag)."""
        hc = _lib.c_vp()
        _lib.check(self.L.lbm_batch_case(self.h, int(c), ctypes.byref(hc)), "lbm_batch_case")
        d = FusedDomain.__new__(FusedDomain)
        d.L, d.h, d.nx, d.ny = self.L, hc, int(self.cases[c]["nx"]), int(self.cases[c]["ny"])
        d.storage, d.ring = self.storage, self.ring
        links = self.cases[c].get("links")
        d.n_obs = 0 if links is None or not len(links) else \
            len(self.cases[c].get("obs_offsets", [0, len(links)])) - 1
        d.close = lambda: None
        return d


def cavity_ensemble(n_cases=256, seed=0, n_lo=64, n_hi=512, re_lo=100.0, re_hi=1000.0,
                    u_lid=0.1):
    """Seeded lid-driven cavities: nx = ny = n ~ U{n_lo..n_hi}, Re ~ logU[re_lo, re_hi],
    parameters as the reference cavity app derives them (cavity.py:36-48)."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(n_lo, n_hi + 1, n_cases)
    res = np.exp(rng.uniform(math.log(re_lo), math.log(re_hi), n_cases))
    cs2 = (1.0 / math.sqrt(3.0)) ** 2
    cases = []
    for n, re in zip(ns, res):
        n = int(n)
        nu = u_lid * n / float(re)
        tau = 0.5 + nu / cs2
        tau_m = 0.25 / (tau - 0.5) + 0.5
        top = np.zeros((2, n))
        top[0, :] = 1.0
        cases.append(dict(nx=n, ny=n, om_p=1.0 / tau, om_m=1.0 / tau_m, Re=float(re), tau=tau,
                          bc=[_lib.BC_VELOCITY] * 4, profiles=[None, None, top, None, None],
                          sigma=math.floor(10 * n), u_lid=u_lid))
    return cases


def cavity_amplitudes(cases, it0, n):
    """Lid amplitude u_lid * (1 - exp(-it^2 / (2 sigma^2))) per case (cavity.py:65-76)."""
    amp = np.zeros((n, len(cases), 5))
    for k in range(n):
        it = it0 + k
        for c, cs in enumerate(cases):
            ret = (1.0 - math.exp(-it ** 2 / (2.0 * cs["sigma"] ** 2)))
            amp[k, c, 2] = cs["u_lid"] * ret
    return amp


def autotune(cases, storage="f32", candidates=None, steps=20, cache=None, log=None):
    """Time each launch shape (cells per work-item, vector width, grid) on the
    actual batch and return the fastest as kwargs for Batch; results are
    cached per (device, storage, case shapes) in a JSON file."""
    import torch
    key = "%s|%s|%d|%d" % (torch.cuda.get_device_name(0), storage, len(cases),
                           hash(tuple((c["nx"], c["ny"]) for c in cases)) & 0xffffffff)
    cache = cache or os.path.join(os.path.expanduser("~"), ".cache", "lbm_amd", "autotune.json")
    try:
        with open(cache) as f:
            known = json.load(f)
    except (OSError, ValueError):
        known = {}
    if key in known:
        return known[key], []
    if candidates is None:
        cus = torch.cuda.get_device_properties(0).multi_processor_count
        cpts = [1, 2, 4] if storage in ("f64", "f32") else [1, 2, 4]
        vws = [1] if storage in ("f64", "f32") else [1, 4]
        candidates = [dict(cells_per_thread=c, vector_width=v, grid=g)
                      for c in cpts for v in vws for g in (0, cus * 8, cus * 16)]
    results = []
    amp = cavity_amplitudes(cases, 0, steps + 2) if "sigma" in cases[0] else \
        np.zeros((steps + 2, len(cases), 5))
    for cand in candidates:
        b = Batch(cases, storage=storage, ring=steps + 4, **cand)
        b.set_amplitudes(0, amp)
        b.init()
        b.steps(1, steps)          # first replay captures the graph
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        b.steps(1, steps)
        e.record()
        torch.cuda.synchronize()
        ms = s.elapsed_time(e) / steps
        eff = dict(cand, cells_per_thread=b.cpt, vector_width=b.vector_width, grid=b.grid)
        results.append((ms, eff))
        if log:
            log(f"# autotune {eff}: {ms * 1e3:.1f} us/iteration")
        b.close()
        del b
    best = min(results, key=lambda r: r[0])[1]
    known[key] = best
    try:
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        with open(cache, "w") as f:
            json.dump(known, f)
    except OSError:
        pass
    return best, results
