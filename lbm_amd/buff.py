"""Observable history with a triple moving average and a convergence flag.

Restates lbm/src/utils/buff.py:9-77 with the same values: ``add`` appends
(amortised O(1) storage instead of ``np.append``), ``mv_avg`` averages the last
fifth of the history three times over and counts steps whose growth stays
under ``obs_cv_ct``; more than ``obs_cv_nb`` such steps in a row set ``obs_cv``.
"""
import math

import numpy as np


class _Grow:
    """Append-only float64 vector, initially [0, 0] as in the reference."""

    def __init__(self):
        self._a = np.zeros(1024)
        self.n = 2

    def reserve(self, extra):
        if self.n + extra > len(self._a):
            cap = len(self._a)
            while cap < self.n + extra:
                cap *= 2
            self._a = np.concatenate([self._a, np.zeros(cap - len(self._a))])

    def append(self, v):
        if self.n == len(self._a):
            self.reserve(1)
        self._a[self.n] = v
        self.n += 1

    @property
    def view(self):
        return self._a[:self.n]


class buff:
    def __init__(self, name, dt, obs_cv_ct, obs_cv_nb, output_dir):
        self.name = name
        self._buff = _Grow()
        self._avg1 = _Grow()
        self._avg2 = _Grow()
        self._avg3 = _Grow()
        self.it = 0
        self.dt = dt
        self.output_dir = output_dir
        self.freq = 1.0
        self.obs = 0.0
        self.obs_cv_ct = obs_cv_ct
        self.obs_cv_nb = obs_cv_nb
        self.obs_cv_cnt = 0
        self.obs_cv = False
        self.obs_cv_it = None     # the value of `it` when obs_cv was set

    @property
    def buff(self):
        return self._buff.view

    @property
    def avg1_buff(self):
        return self._avg1.view

    @property
    def avg2_buff(self):
        return self._avg2.view

    @property
    def avg3_buff(self):
        return self._avg3.view

    def add(self, value):
        self._buff.append(value)
        self.it += 1

    def f_avg(self):
        b = self.buff
        return np.sum(b) / len(b)

    def p_avg(self, b, i, j):
        return np.sum(b[i:j]) / (float(j - i + 1))

    def add_chunk(self, values):
        """``add(v); mv_avg()`` for each v of ``values`` in order, with the
        same results (the window sums run in the D2Q9 library's host code,
        lbm_mv_avg_chunk, bit for bit what np.sum gives).  Returns the arrays
        of what mv_avg returned: (obs, growth)."""
        from lbm_amd import _lib
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        k = len(v)
        obs = np.zeros(k)
        growth = np.zeros(k)
        if k == 0:
            return obs, growth
        for g in (self._buff, self._avg1, self._avg2, self._avg3):
            g.reserve(k)
        it0 = self.it
        n0 = self._buff.n
        self._buff._a[n0:n0 + k] = v
        self._buff.n += k
        self.it += k
        ptr = lambda g: g._a.ctypes.data_as(_lib.c_vp)
        _lib.check(_lib.load().lbm_mv_avg_chunk(ptr(self._buff), ptr(self._avg1), ptr(self._avg2),
                                                ptr(self._avg3), it0, k,
                                                obs.ctypes.data_as(_lib.c_vp)),
                   "lbm_mv_avg_chunk")
        for g in (self._avg1, self._avg2, self._avg3):
            g.n += k
        a3 = self._avg3._a
        self._chunk = dict(k=k, obs=(self.obs, self.obs_cv_cnt, self.obs_cv, self.obs_cv_it),
                           cnt=np.zeros(k, dtype=np.int64), obs_vals=obs)
        for t in range(k):
            it = it0 + 1 + t
            if it <= 10:
                continue
            it_s = math.floor(it * 4 / 5)
            db = a3[it] - a3[it_s]
            gr = db / ((it - it_s + 1) * self.dt)
            growth[t] = gr
            if abs(gr) < self.obs_cv_ct:
                self.obs_cv_cnt += 1
            else:
                self.obs_cv_cnt = 0
            if self.obs_cv_cnt > self.obs_cv_nb and not self.obs_cv:
                self.obs_cv = True
                self.obs_cv_it = it
            self._chunk["cnt"][t] = self.obs_cv_cnt
        self.obs = obs[-1]
        return obs, growth

    def truncate_chunk(self, m):
        """Keep only the first m values of the last add_chunk call: the
        history, averages, counters and flag as after those m iterations."""
        c = self._chunk
        drop = c["k"] - m
        if drop <= 0:
            return
        for g in (self._buff, self._avg1, self._avg2, self._avg3):
            g.n -= drop
            g._a[g.n:g.n + drop] = 0.0
        self.it -= drop
        obs0, cnt0, cv0, cvit0 = c["obs"]
        if m == 0:
            self.obs, self.obs_cv_cnt, self.obs_cv, self.obs_cv_it = obs0, cnt0, cv0, cvit0
        else:
            self.obs = c["obs_vals"][m - 1]
            self.obs_cv_cnt = int(c["cnt"][m - 1])
            if self.obs_cv and not cv0 and self.obs_cv_it > self.it:
                self.obs_cv, self.obs_cv_it = False, None
        c["k"] = m

    def mv_avg(self):
        it_s = math.floor(self.it * 4 / 5)
        it_e = self.it
        self.obs = self.p_avg(self.buff, it_s, it_e)
        self._avg1.append(self.obs)
        self.obs = self.p_avg(self.avg1_buff, it_s, it_e)
        self._avg2.append(self.obs)
        self.obs = self.p_avg(self.avg2_buff, it_s, it_e)
        self._avg3.append(self.obs)
        self.obs = self.p_avg(self.avg3_buff, it_s, it_e)
        growth = 0.0
        if self.it > 10:
            a3 = self.avg3_buff
            db = a3[it_e] - a3[it_s]
            growth = db / ((it_e - it_s + 1) * self.dt)
            if abs(growth) < self.obs_cv_ct:
                self.obs_cv_cnt += 1
            else:
                self.obs_cv_cnt = 0
            if self.obs_cv_cnt > self.obs_cv_nb and not self.obs_cv:
                self.obs_cv = True
                self.obs_cv_it = self.it
        return self.obs, growth
