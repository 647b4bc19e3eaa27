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


def ref_cavity(precision="f64"):
    """The reference's default cavity app (100x100, Re = 100, u_lid = 0.2,
    it_max = 10 000), as lbm/tst/cavity/test_cavity.py runs it."""
    app = cavity()
    app.__class__ = type("ref_cavity", (_Quiet, cavity), {})
    app.precision = precision
    return app


def ref_poiseuille(precision="f64"):
    """The reference's default Poiseuille app (150x50, it_max = 11 250), as
    lbm/tst/poiseuille/test_poiseuille.py runs it."""
    from lbm_amd.app.poiseuille import poiseuille
    app = poiseuille()
    app.__class__ = type("ref_poiseuille", (_Quiet, poiseuille), {})
    app.precision = precision
    return app


def cavity200_app(precision="f64"):
    """Cavity 200x200, u_lid = 0.1, Re = 100, 20 001 iterations."""
    from lbm_amd.app.configs import cavity200
    app = cavity200(precision)
    app.__class__ = type("q_cavity200", (_Quiet, cavity200), {})
    return app


def turek200_app(precision="f64"):
    """Schaefer-Turek channel 1073x200 (IBB cylinder D = 48, u_max = 0.05,
    Re = 100, sigma = 10730), Cd/Cl recorded every iteration."""
    from lbm_amd.app.configs import turek200
    app = turek200(precision)
    app.__class__ = type("q_turek200", (_Quiet, _RecordDrag, turek200), {})
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


def array_domain_case(storage, iterations=80, vector_width=0, cpt=0, ny=64, n_obs=3, seed=1,
                      side=(6, 14), aspect=4):
    """Obstacle-array channel (nx = aspect ny) run directly through FusedDomain;
    returns the final fields (epilogue), the drag series and the domain shape.
    side: range of the squares' side lengths."""
    import math
    import types

    import torch

    from lbm_amd import _lib, geometry
    from lbm_amd.app.configs import obstacle_array
    from lbm_amd.domain import FusedDomain, pack_inlets
    from lbm_amd.shapes import generate_shape
    app = obstacle_array(ny=ny, aspect=aspect, n_obs=n_obs, seed=seed, precision=storage,
                         iterations=iterations, sigma=20, n_spts=100, side_lo=side[0],
                         side_hi=side[1])
    lat = types.SimpleNamespace(nx=app.nx, ny=app.ny, x_min=app.x_min, x_max=app.x_max,
                                y_min=app.y_min, y_max=app.y_max, dx=app.dx, IBB=True,
                                lattice=np.zeros((app.nx, app.ny)))
    for k, o in enumerate(app.obstacles):
        s = generate_shape(o.n_pts, o.pos, o.type, o.size, o.name, o.n_spts, _TMP + "/")
        o.set_polygon(s.curve_pts)
        o.set_tag(k + 1)
        o.fill(*geometry.add_obstacle(lat, o, verbose=False))
    tau_p = app.tau_lbm
    om_p, om_m = 1.0 / tau_p, 1.0 / (0.25 / (tau_p - 0.5) + 0.5)
    nx, ny = app.nx, app.ny
    bc = [_lib.BC_VELOCITY, _lib.BC_VELOCITY, _lib.BC_VELOCITY, _lib.BC_PRESSURE]
    dom = FusedDomain(nx, ny, om_p, om_m, bc, [(o.boundary, o.ibb) for o in app.obstacles],
                      ibb=True, storage=storage, ring=iterations + 2, cells_per_thread=cpt,
                      vector_width=vector_width)
    rows = []
    for it in range(iterations + 1):
        ret = (1.0 - math.exp(-it ** 2 / (2.0 * app.sigma ** 2)))
        ul = np.zeros((2, ny))
        ul[0, :] = ret * app.u_lbm
        rows.append((np.zeros((2, nx)), ul, np.zeros((2, nx)), np.zeros((2, ny)),
                     np.full(ny, 1.0)))
    pool, offs = pack_inlets(rows)
    dom.set_inlets(0, offs, pool)
    w = torch.tensor([4. / 9.] + [1. / 9.] * 4 + [1. / 36.] * 4, dtype=torch.float64)
    g0 = w[:, None, None].expand(9, nx, ny).contiguous().cuda()
    rho = torch.empty((nx, ny), dtype=torch.float64, device="cuda")
    u = torch.empty((2, nx, ny), dtype=torch.float64, device="cuda")
    dom.init(g0)
    for it in range(1, iterations - 1):
        dom.step(it)
    last = iterations - 1
    dom.step(last, rho, u)
    g = torch.empty((9, nx, ny), dtype=torch.float64, device="cuda")
    gup = torch.empty_like(g)
    dom.drag_only(last)
    drag = dom.drag_read(0, iterations)
    dom.epilogue(last, g, gup, rho, u)
    out = dict(rho=rho.cpu().numpy(), u=u.cpu().numpy(), g=g.cpu().numpy(),
               g_up=gup.cpu().numpy(), drag=drag, app=app, om=(om_p, om_m), rows=rows)
    dom.close()
    return out


def array_oracle_case(gpu_out, iterations=80):
    """The same run on the oracle, from the GPU case's geometry and inlets."""
    from oracle import oracle
    app = gpu_out["app"]
    om_p, om_m = gpu_out["om"]
    nx, ny = app.nx, app.ny
    w = np.array([4. / 9.] + [1. / 9.] * 4 + [1. / 36.] * 4)
    g = np.ascontiguousarray(np.broadcast_to(w[:, None, None], (9, nx, ny)))
    geq, gup = np.zeros_like(g), np.zeros_like(g)
    rho, u = np.ones((nx, ny)), np.zeros((2, nx, ny))
    drag = []
    for it in range(iterations):
        ub, ul, ut, ur, rr = gpu_out["rows"][it]
        oracle.macro(g, rho, u)
        oracle.equilibrium(u, rho, geq)
        oracle.col_str(g, geq, gup, om_p, om_m)
        for o in app.obstacles:
            oracle.bounce_back_obstacle(True, o.boundary, o.ibb, gup, g)
        oracle.zou_he("bottom", u, rho, g, side=np.ascontiguousarray(ub))
        oracle.zou_he("left", u, rho, g, side=np.ascontiguousarray(ul))
        oracle.zou_he("top", u, rho, g, side=np.ascontiguousarray(ut))
        oracle.zou_he("right_pressure", u, rho, g, side=np.ascontiguousarray(ur),
                      rho_right=np.ascontiguousarray(rr))
        for c in ("bottom_left", "top_left", "top_right", "bottom_right"):
            oracle.zou_he(c, u, rho, g)
        drag.append([oracle.drag_sums(o.boundary, gup, g) for o in app.obstacles])
    return dict(rho=rho, u=u, g=g, g_up=gup, drag=np.array(drag))


def batch_case_oracle(cs, amp_c, iterations):
    """One case of a batched ensemble (a Batch case dict) run alone on the
    oracle: rest state at iteration 0, wall arrays amplitude(it, side) x
    profile(side), the reference loop of run.py:24-54 with the set_bc order
    of the reference app of that kind (cavity.py:79-89: bottom, left, right,
    top; turek.py:112-125: obstacles, bottom, left, top, right), drag sums
    of every obstacle and iteration."""
    from lbm_amd import _lib
    from oracle import oracle
    nx, ny = int(cs["nx"]), int(cs["ny"])
    w = np.array([4. / 9.] + [1. / 9.] * 4 + [1. / 36.] * 4)
    g = np.ascontiguousarray(np.broadcast_to(w[:, None, None], (9, nx, ny)))
    geq, gup = np.zeros_like(g), np.zeros_like(g)
    rho, u = np.ones((nx, ny)), np.zeros((2, nx, ny))
    shapes = ((2, nx), (2, ny), (2, nx), (2, ny), (ny,))
    profs = cs.get("profiles", [None] * 5)
    links = cs.get("links")
    obs = []
    if links is not None and len(links):
        off = list(cs.get("obs_offsets", [0, len(links)]))
        ibb_on = np.asarray(cs.get("ibb_on", np.zeros(len(links))))
        assert ibb_on.min() == ibb_on.max(), "one IBB flag per case"
        ibb = bool(ibb_on[0])
        for a, b in zip(off[:-1], off[1:]):
            obs.append((np.asarray(links[a:b], np.int64),
                        np.asarray(cs.get("ibb", np.zeros(len(links)))[a:b], np.float64)))
    pressure = cs["bc"][3] == _lib.BC_PRESSURE
    drag = []
    for it in range(iterations):
        side = [np.zeros(s) if p is None else amp_c[it, sd] * np.asarray(p, np.float64)
                for sd, (s, p) in enumerate(zip(shapes, profs))]
        side = [np.ascontiguousarray(a) for a in side]
        oracle.macro(g, rho, u)
        oracle.equilibrium(u, rho, geq)
        oracle.col_str(g, geq, gup, cs["om_p"], cs["om_m"])
        for bnd, dist in obs:
            oracle.bounce_back_obstacle(ibb, bnd, dist, gup, g)
        oracle.zou_he("bottom", u, rho, g, side=side[0])
        oracle.zou_he("left", u, rho, g, side=side[1])
        if pressure:
            oracle.zou_he("top", u, rho, g, side=side[2])
            oracle.zou_he("right_pressure", u, rho, g, side=side[3], rho_right=side[4])
        else:
            oracle.zou_he("right", u, rho, g, side=side[3])
            oracle.zou_he("top", u, rho, g, side=side[2])
        for c in ("bottom_left", "top_left", "top_right", "bottom_right"):
            oracle.zou_he(c, u, rho, g)
        drag.append([oracle.drag_sums(bnd, gup, g) for bnd, _ in obs])
    return dict(rho=rho, u=u, g=g, g_up=gup, drag=np.array(drag) if obs else None)


def one_step_pair(storage="f16", cpt=1, ny=64, nx=256, seed=5, iterations=2, shapes=None):
    """Two FusedDomains of the same channel (no obstacles), advanced from one
    random state; returns the g_up of the last iteration of each (decoded to
    double).  shapes: the two (cells_per_thread, vector_width) launch shapes,
    by default the 4-node and 1-node 16-bit kernels with cpt nodes."""
    import torch
    from lbm_amd import _lib
    from lbm_amd.domain import FusedDomain, pack_inlets
    rng = np.random.default_rng(seed)
    w = np.array([4. / 9.] + [1. / 9.] * 4 + [1. / 36.] * 4)
    g0 = w[:, None, None] * (1.0 + 0.05 * rng.standard_normal((9, nx, ny)))
    g0 = torch.from_numpy(np.ascontiguousarray(g0)).cuda()
    bc = [_lib.BC_VELOCITY, _lib.BC_VELOCITY, _lib.BC_VELOCITY, _lib.BC_PRESSURE]
    rows = [(np.zeros((2, nx)), np.zeros((2, ny)), np.zeros((2, nx)), np.zeros((2, ny)),
             np.ones(ny))] * (iterations + 1)
    pool, offs = pack_inlets(rows)
    outs = []
    for c, vw in (shapes or ((cpt, 4), (cpt, 1))):
        dom = FusedDomain(nx, ny, 1 / 0.6, 1 / (0.25 / 0.1 + 0.5), bc, storage=storage,
                          ring=iterations + 2, cells_per_thread=c, vector_width=vw)
        dom.set_inlets(0, offs, pool)
        dom.init(g0)
        for it in range(1, iterations + 1):
            dom.step(it)
        g = torch.empty((9, nx, ny), dtype=torch.float64, device="cuda")
        gup = torch.empty_like(g)
        dom.epilogue(iterations, g, gup)
        outs.append(gup.cpu().numpy())
        dom.close()
    return outs
