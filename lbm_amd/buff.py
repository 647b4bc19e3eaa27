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

    def append(self, v):
        if self.n == len(self._a):
            self._a = np.concatenate([self._a, np.zeros(len(self._a))])
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
            if self.obs_cv_cnt > self.obs_cv_nb:
                self.obs_cv = True
        return self.obs, growth
