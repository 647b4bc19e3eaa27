"""The Turek app's chunk replay (replay_chunk: vectorised drag/lift, moving
averages and drag_lift lines) against its per-iteration hooks (printings,
outputs, observables, check_stop in the order of lbm/src/core/run.py:24-54)
on the same drag/lift sums: identical drag_lift file, histories, averages,
convergence state and stopping iteration, for stop = 'it' and stop = 'obs'
with the stop inside a chunk.  Host code only (the sums are synthetic)."""
import os
import tempfile

import numpy as np
import pytest

from lbm_amd.app.configs import turek200
from lbm_amd.app.turek import turek
from lbm_amd.buff import buff
from lbm_amd.lattice import lattice


class _Replay:
    def __init__(self, fx, fy, obs):
        self.fx, self.fy, self.obs = fx, fy, obs
        self.it = self.it0 = 0
        self.it1 = 0

    def drag_sums(self, obs):
        assert obs is self.obs
        return float(self.fx[self.it]), float(self.fy[self.it])

    def drag_sums_chunk(self, obs):
        assert obs is self.obs
        return self.fx[self.it0:self.it1].copy(), self.fy[self.it0:self.it1].copy()


def _app(kind, tmp, tag):
    app = turek() if kind == "obs" else turek200(iterations=2600)
    app.verbose = False
    app.write_png = False
    lat = lattice(app)
    app.drag_buff = buff("drag", 0.05, 2e-3, 150, tmp)
    app.lift_buff = buff("lift", 0.05, 2e-3, 150, tmp)
    app._dl_file = open(os.path.join(tmp, "drag_lift_" + tag), "w")
    app.output_it = 0
    return app, lat


def _sums(n, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(n)
    fx = -0.08 * (1 - np.exp(-t / 200.0)) * (1 + 0.01 * np.sin(t / 30.0)) \
        - 1e-6 * rng.standard_normal(n)
    fy = 0.02 * np.exp(-t / 400.0) * np.sin(t / 25.0) + 1e-7 * rng.standard_normal(n)
    return fx, fy


@pytest.mark.parametrize("kind", ["it", "obs"])
def test_replay_chunk_equals_per_iteration_hooks(kind):
    n = 6000
    fx, fy = _sums(n, 3)
    with tempfile.TemporaryDirectory() as tmp:
        a, la = _app(kind, tmp, "a")
        b, lb = _app(kind, tmp, "b")
        # per-iteration hooks
        ra = _Replay(fx, fy, a.obstacles[0])
        la._replay = ra
        stop_a = None
        for it in range(n):
            ra.it = it
            a.printings(it)
            a.outputs(la, it)
            a.observables(la, it)
            if not a.check_stop(it):
                stop_a = it
                break
        la._replay = None
        # chunks of 500 iterations (output_freq)
        rb = _Replay(fx, fy, b.obstacles[0])
        lb._replay = rb
        stop_b = None
        for it0 in range(0, n, 500):
            rb.it0, rb.it1 = it0, min(it0 + 500, n)

            def at(it):
                rb.it = it
            stop_b = b.replay_chunk(lb, it0, rb.it1, at)
            if stop_b is not None:
                break
        lb._replay = None
        a._dl_file.close()
        b._dl_file.close()
        assert stop_a is not None and stop_a == stop_b and stop_a % 500 not in (0, 499)
        text_a = open(os.path.join(tmp, "drag_lift_a")).read()
        text_b = open(os.path.join(tmp, "drag_lift_b")).read()
        lines_a, lines_b = text_a.splitlines(), text_b.splitlines()
        assert len(lines_a) == len(lines_b) == stop_a + 1
        bad = next((i for i, (x, y) in enumerate(zip(lines_a, lines_b)) if x != y), None)
        assert bad is None, (bad, lines_a[bad], lines_b[bad])
        for name in ("drag_buff", "lift_buff"):
            x, y = getattr(a, name), getattr(b, name)
            for arr in ("buff", "avg1_buff", "avg2_buff", "avg3_buff"):
                assert np.array_equal(getattr(x, arr), getattr(y, arr)), (name, arr)
            assert (x.it, x.obs, x.obs_cv_cnt, x.obs_cv, x.obs_cv_it) == \
                (y.it, y.obs, y.obs_cv_cnt, y.obs_cv, y.obs_cv_it)


def test_replay_chunk_declines_overridden_hooks():
    class sub(turek):
        def observables(self, lattice, it):
            pass
    app = sub()
    assert app.replay_chunk(None, 0, 10, lambda it: None) is NotImplemented
