"""CPU oracle of the D2Q9 core -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  The product (lbm_amd) never does.

``OracleLattice`` restates the reference lattice (lbm/src/core/lattice.py) on
NumPy arrays; its hot-path methods call the strict-order C restatement of
lbm/src/core/nb.py in d2q9_oracle.c (built by oracle/Makefile into
oracle/_build/).  ``run`` restates the reference time loop (lbm/src/core/run.py).

Pins of this oracle (each one a check a plausible mistake would fail):
  * reference golden data: lbm/tst/cavity/test_cavity.py's four asserted
    values (1e-6) and all 200 values of cavity_{vx,uy}_ref (5e-6; the files
    were written by the fastmath Numba/BLAS build, measured deviation 3.2e-6),
    100x100 cavity, 10 001 steps -- covers macro, equilibrium, TRT collision,
    streaming, all four Zou-He velocity walls and the corners;
  * reference test_poiseuille (L1 < 1e-3 against the analytic parabola) --
    covers the parabolic Zou-He inlet and the Zou-He pressure outlet;
  * closed forms and invariants (tests/test_oracle.py): equilibrium moments,
    collision invariance of rho and rho*u, streaming along c_q, the rest
    state as a fixed point, Zou-He walls reproducing the imposed u or
    density, corners at rest, drag of a single link and of a symmetric body;
  * IBB (tests/test_oracle_ibb.py): exactness of the quadratic
    Bouzidi-Firdaouss-Lallemand interpolation on random quadratics for
    p in {0.1, 0.3, 0.5, 0.7, 0.95} and all eight directions, with unrelated
    values around the stencil, plus planted plane / offset / direction / sign
    mistakes in copies of d2q9_oracle.c that this check must catch.
Every function has such a check; none is "parity unpinned".
"""
import ctypes
import math
import os
import subprocess
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle_d2q9.so")

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    i = ctypes.c_int
    d = ctypes.c_double
    sig = {
        "oracle_macro": [i, i, vp, vp, vp],
        "oracle_equilibrium": [i, i, vp, vp, vp],
        "oracle_col_str": [i, i, vp, vp, vp, d, d],
        "oracle_drag_sums": [i, i, i, vp, vp, vp, vp],
        "oracle_bounce_back_obstacle": [i, i, i, i, vp, vp, vp, vp],
        "oracle_zou_he_left_wall_velocity": [i, i, vp, vp, vp, vp],
        "oracle_zou_he_right_wall_velocity": [i, i, vp, vp, vp, vp],
        "oracle_zou_he_right_wall_pressure": [i, i, vp, vp, vp, vp, vp],
        "oracle_zou_he_top_wall_velocity": [i, i, vp, vp, vp, vp],
        "oracle_zou_he_bottom_wall_velocity": [i, i, vp, vp, vp, vp],
        "oracle_zou_he_bottom_left_corner": [i, i, vp, vp, vp],
        "oracle_zou_he_top_left_corner": [i, i, vp, vp, vp],
        "oracle_zou_he_top_right_corner": [i, i, vp, vp, vp],
        "oracle_zou_he_bottom_right_corner": [i, i, vp, vp, vp],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = i if name in ("oracle_drag_sums", "oracle_bounce_back_obstacle") else None
    lib.oracle_threads.argtypes = []
    lib.oracle_threads.restype = i
    _lib = lib
    return lib


def threads():
    """Threads of the oracle's node loops (OpenMP, OMP_NUM_THREADS)."""
    return int(load().oracle_threads())


def _p(a):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return _p(a)


# ---------------------------------------------------------------- kernels --
def macro(g, rho, u):
    """lattice.macro, lattice.py:178-189."""
    nx, ny = rho.shape
    load().oracle_macro(nx, ny, _f64(g), _f64(rho), _f64(u))


def equilibrium(u, rho, g_eq):
    """nb_equilibrium, nb.py:7-17."""
    nx, ny = rho.shape
    load().oracle_equilibrium(nx, ny, _f64(u), _f64(rho), _f64(g_eq))


def col_str(g, g_eq, g_up, om_p, om_m):
    """nb_col_str, nb.py:21-45."""
    _, nx, ny = g.shape
    load().oracle_col_str(nx, ny, _f64(g), _f64(g_eq), _f64(g_up), om_p, om_m)


def drag_sums(boundary, g_up, g):
    """nb_drag_lift sums (fx, fy), nb.py:49-67."""
    _, nx, ny = g.shape
    b = np.ascontiguousarray(boundary, dtype=np.int64)
    out = np.zeros(2)
    st = load().oracle_drag_sums(nx, ny, b.shape[0], _p(b), _f64(g_up), _f64(g), _f64(out))
    if st != 0:
        raise IndexError("boundary index out of range")
    return float(out[0]), float(out[1])


def drag_lift(boundary, g_up, g, R_ref, U_ref, L_ref):
    """nb_drag_lift, nb.py:49-73 (normalisation as written at :70-71)."""
    fx, fy = drag_sums(boundary, g_up, g)
    Cx = -2.0 * fx / (R_ref * L_ref * U_ref ** 2)
    Cy = -2.0 * fy / (R_ref * L_ref * U_ref ** 2)
    return Cx, Cy


def bounce_back_obstacle(IBB, boundary, obs_ibb, g_up, g):
    """nb_bounce_back_obstacle, nb.py:77-117."""
    _, nx, ny = g.shape
    b = np.ascontiguousarray(boundary, dtype=np.int64)
    ib = np.ascontiguousarray(obs_ibb, dtype=np.float64) if IBB else np.zeros(max(len(b), 1))
    st = load().oracle_bounce_back_obstacle(nx, ny, int(bool(IBB)), b.shape[0], _p(b),
                                            _f64(ib), _f64(g_up), _f64(g))
    if st != 0:
        raise IndexError("boundary or IBB stencil index out of range")


def zou_he(kind, u, rho, g, side=None, rho_right=None):
    """The nine nb_zou_he_* routines, nb.py:121-344."""
    _, nx, ny = g.shape
    L = load()
    if kind == "left":
        L.oracle_zou_he_left_wall_velocity(nx, ny, _f64(u), _f64(side), _f64(rho), _f64(g))
    elif kind == "right":
        L.oracle_zou_he_right_wall_velocity(nx, ny, _f64(u), _f64(side), _f64(rho), _f64(g))
    elif kind == "right_pressure":
        L.oracle_zou_he_right_wall_pressure(nx, ny, _f64(u), _f64(rho_right), _f64(side),
                                            _f64(rho), _f64(g))
    elif kind == "top":
        L.oracle_zou_he_top_wall_velocity(nx, ny, _f64(u), _f64(side), _f64(rho), _f64(g))
    elif kind == "bottom":
        L.oracle_zou_he_bottom_wall_velocity(nx, ny, _f64(u), _f64(side), _f64(rho), _f64(g))
    elif kind in ("bottom_left", "top_left", "top_right", "bottom_right"):
        getattr(L, f"oracle_zou_he_{kind}_corner")(nx, ny, _f64(u), _f64(rho), _f64(g))
    else:
        raise ValueError(kind)


# ---------------------------------------------------------------- lattice --
class OracleLattice:
    """Restatement of the reference ``lattice`` (lattice.py:15-436) on NumPy
    arrays, with the hot-path methods calling the C oracle."""

    def __init__(self, app, results_dir="./results/"):
        # lattice.py:20-78 (defaults, then app attributes)
        d = dict(name="lattice", x_min=0.0, x_max=1.0, y_min=0.0, y_max=1.0, nx=100, ny=100,
                 tau_lbm=1.0, dx=1.0, dt=1.0, dpi=100, u_lbm=0.03, L_lbm=100, nu_lbm=0.01,
                 Re_lbm=100.0, rho_lbm=1.0, IBB=False, stop="it", t_max=1.0, it_max=1000,
                 obs_cv_ct=1.0e-1, obs_cv_nb=500)
        for k, v in d.items():
            setattr(self, k, getattr(app, k, v))
        stamp = time.strftime("%Y-%m-%d_%H_%M_%S")
        self.results_dir = results_dir
        self.output_dir = os.path.join(results_dir, "oracle_" + stamp) + "/"
        self.png_dir = self.output_dir + "./png/"
        os.makedirs(self.png_dir, exist_ok=True)
        self.set_default_lbm()

    def set_default_lbm(self):
        """lattice.py:117-174."""
        self.output_it = 0
        self.lx = self.nx - 1
        self.ly = self.ny - 1
        self.q = 9
        self.Cs = 1.0 / math.sqrt(3.0)
        self.tau_p_lbm = self.tau_lbm
        self.lambda_trt = 1.0 / 4.0
        self.tau_m_lbm = self.lambda_trt / (self.tau_p_lbm - 0.5) + 0.5
        self.om_p_lbm = 1.0 / self.tau_p_lbm
        self.om_m_lbm = 1.0 / self.tau_m_lbm
        self.om_lbm = 1.0 / self.tau_lbm
        self.c = np.array([[0, 0], [1, 0], [-1, 0], [0, 1], [0, -1],
                           [1, 1], [-1, -1], [-1, 1], [1, -1]])
        self.w = np.array([4. / 9.] + [1. / 9.] * 4 + [1. / 36.] * 4)
        self.ns = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7])
        self.g = np.zeros((9, self.nx, self.ny))
        self.g_eq = np.zeros((9, self.nx, self.ny))
        self.g_up = np.zeros((9, self.nx, self.ny))
        self.u_left = np.zeros((2, self.ny))
        self.u_right = np.zeros((2, self.ny))
        self.u_top = np.zeros((2, self.nx))
        self.u_bot = np.zeros((2, self.nx))
        self.rho_right = np.zeros(self.ny)
        self.lattice = np.zeros((self.nx, self.ny))
        self.rho = np.ones((self.nx, self.ny))
        self.u = np.zeros((2, self.nx, self.ny))

    # hot path
    def macro(self):
        macro(self.g, self.rho, self.u)

    def equilibrium(self):
        equilibrium(self.u, self.rho, self.g_eq)

    def collision_stream(self):
        col_str(self.g, self.g_eq, self.g_up, self.om_p_lbm, self.om_m_lbm)

    def drag_lift(self, obs, R_ref, U_ref, L_ref):
        return drag_lift(obs.boundary, self.g_up, self.g, R_ref, U_ref, L_ref)

    def bounce_back_obstacle(self, obstacle):
        bounce_back_obstacle(self.IBB, obstacle.boundary, obstacle.ibb, self.g_up, self.g)

    def zou_he_left_wall_velocity(self):
        zou_he("left", self.u, self.rho, self.g, side=self.u_left)

    def zou_he_right_wall_velocity(self):
        zou_he("right", self.u, self.rho, self.g, side=self.u_right)

    def zou_he_right_wall_pressure(self):
        zou_he("right_pressure", self.u, self.rho, self.g, side=self.u_right,
               rho_right=self.rho_right)

    def zou_he_top_wall_velocity(self):
        zou_he("top", self.u, self.rho, self.g, side=self.u_top)

    def zou_he_bottom_wall_velocity(self):
        zou_he("bottom", self.u, self.rho, self.g, side=self.u_bot)

    def zou_he_bottom_left_corner(self):
        zou_he("bottom_left", self.u, self.rho, self.g)

    def zou_he_top_left_corner(self):
        zou_he("top_left", self.u, self.rho, self.g)

    def zou_he_top_right_corner(self):
        zou_he("top_right", self.u, self.rho, self.g)

    def zou_he_bottom_right_corner(self):
        zou_he("bottom_right", self.u, self.rho, self.g)

    # setup (literal restatement, oracle/geometry.py)
    def add_obstacle(self, obstacle):
        from oracle import geometry
        return geometry.add_obstacle(self, obstacle)

    def get_coords(self, i, j):
        from oracle import geometry
        return geometry.get_coords(self, i, j)

    def generate_image(self, obstacles):
        pass


def run(lattice, app, on_step=None):
    """The reference time loop, run.py:12-61 (without the wall-clock print)."""
    app.initialize(lattice)
    it = 0
    compute = True
    while compute:
        app.printings(it)
        app.set_inlets(lattice, it)
        lattice.macro()
        app.outputs(lattice, it)
        lattice.equilibrium()
        lattice.collision_stream()
        app.set_bc(lattice)
        app.observables(lattice, it)
        if on_step is not None:
            on_step(lattice, it)
        compute = app.check_stop(it)
        it += 1
    app.finalize(lattice)
    return it
