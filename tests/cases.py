"""Small cases run through both the fused GPU path (lbm_amd.lattice + run)
and the CPU oracle (oracle.OracleLattice + oracle.run) on identical inputs."""
import math
import tempfile

import numpy as np

from lbm_amd.app.cavity import cavity
from lbm_amd.app.turek import turek
from lbm_amd.app.step import step
from lbm_amd.app.array import array

_TMP = tempfile.mkdtemp(prefix="lbm_amd_tests_")


class _Quiet:
    verbose = False
    write_png = False
    results_dir = _TMP + "/"


def small_cavity(precision="f64"):
    """Cavity 48x40 (non-square), lid ramp sigma = 20 iterations."""
    app = cavity()
    app.__class__ = type("small_cavity", (_Quiet, cavity), {})
    app.L_lbm = 40
    app.x_max = 1.2
    app.u_lbm = 0.1
    app.compute_lbm_parameters()
    app.sigma = 20
    app.output_freq = 25
    app.precision = precision
    return app


class _RecordDrag:
    def initialize(self, lattice):
        super().initialize(lattice)
        self.drag_series = []

    def observables(self, lattice, it):
        self.drag_series.append([lattice.drag_lift(o, self.rho_lbm, self.u_avg, self.D_lbm)
                                 for o in self.obstacles])

    def printings(self, it):
        pass

    def check_stop(self, it):
        stop_at = getattr(self, "stop_at", None)
        if stop_at is not None:
            return it < stop_at
        return it < self.it_max

    def finalize(self, lattice):
        pass


def small_channel(precision="f64", L=40, ibb=True):
    """Turek channel ny=40 (nx=214, IBB cylinder D=9), ramp sigma = 30."""
    app = turek()
    app.__class__ = type("small_channel", (_Quiet, _RecordDrag, turek), {})
    app.L_lbm = L
    app.compute_lbm_parameters()
    app.obstacles[0].size = app.r_cyl
    app.IBB = ibb
    app.sigma = 30
    app.stop = 'it'
    app.output_freq = 25
    app.precision = precision
    return app


def small_step(precision="f64"):
    """The reference step app at ny=30 (plain bounce-back, squares on the walls,
    links wrapped to the top row), sigma = 30."""
    app = step()
    app.__class__ = type("small_step", (_Quiet, _RecordDrag, step), {})
    app.L_lbm = 30
    app.compute_lbm_parameters()
    app.sigma = 30
    app.output_freq = 25
    app.precision = precision
    return app


def small_array(precision="f64"):
    """The reference array app at ny=60 (8 IBB squares), sigma = 30."""
    app = array()
    app.__class__ = type("small_array", (_Quiet, _RecordDrag, array), {})
    app.L_lbm = 60
    app.compute_lbm_parameters()
    app.sigma = 30
    app.output_freq = 25
    app.precision = precision
    return app


def _collect(lat, app):
    out = {k: np.array(getattr(lat, k), copy=True) for k in ("rho", "u", "g", "g_up", "g_eq")}
    out["drag"] = np.array(app.drag_series) if hasattr(app, "drag_series") else None
    return out


def run_gpu(make, iterations, chunk=16, fused=True, **kw):
    from lbm_amd.lattice import lattice
    from lbm_amd.run import run
    app = make(**kw)
    app.it_max = iterations - 1
    lat = lattice(app)
    run(lat, app, chunk=chunk, fused=fused)
    return _collect(lat, app), app, lat


def run_oracle(make, iterations, **kw):
    from oracle import oracle
    app = make(**kw)
    app.it_max = iterations - 1
    lat = oracle.OracleLattice(app, results_dir=_TMP + "/")
    oracle.run(lat, app)
    return _collect(lat, app), app, lat


def run_pair(make, iterations, chunk=16, **kw):
    gpu = run_gpu(make, iterations, chunk=chunk, **kw)[0]
    ora = run_oracle(make, iterations, **kw)[0]
    return gpu, ora
