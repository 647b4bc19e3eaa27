"""Drop-in ``lattice``: the reference's lattice API over HIP kernels on one GPU.

Same constructor, attributes and methods as lbm/src/core/lattice.py:15-436.
The populations g, g_eq, g_up and the fields rho, u live in HBM as float64
tensors with the reference layout ([9][nx][ny], y fastest).  Reading one of
these attributes returns a host NumPy array synchronised at that moment;
assigning one (or mutating the array returned by the read, before the next
lattice call) is uploaded before the next kernel needs it.  The boundary
inputs (u_left, u_right, u_top, u_bot, rho_right) and the obstacle map
``lattice`` are plain host arrays, as in the reference.

Each hot-path method launches the HIP kernel of the reference function it
replaces (include/lbm_d2q9.h); there is no CPU fallback: without a GPU or
without liblbm_d2q9.so these methods raise.  ``lbm_amd.run.run`` drives the
fused time step instead of these per-call kernels.

Extension: an app attribute ``precision`` ('f64' default, 'f32', 'f16',
'bf16', 'f16e') selects the population storage of the fused time step.
"""
import math
import os
from datetime import datetime

import numpy as np

from lbm_amd import _lib
from lbm_amd import geometry


class _DevField:
    """A float64 array held in HBM with a lazily synchronised host mirror."""

    def __init__(self, host):
        self.host = host
        self.dev = None
        self.where = "host"   # "host": host copy is authoritative; "dev": device copy

    def read_host(self):
        if self.where == "dev":
            self.host[...] = self.dev.cpu().numpy()
        self.where = "host"   # the caller may mutate the returned array
        return self.host

    def assign(self, value):
        value = np.asarray(value, dtype=np.float64)
        if value.shape == self.host.shape:
            np.copyto(self.host, value)
        else:
            self.host = np.array(value, dtype=np.float64)
            self.dev = None
        self.where = "host"

    def device(self, load=True):
        import torch
        if self.dev is None or tuple(self.dev.shape) != self.host.shape:
            self.dev = torch.empty(self.host.shape, dtype=torch.float64, device="cuda")
            if not load and self.where == "host":
                self.where = "host"
        if load and self.where == "host":
            self.dev.copy_(torch.from_numpy(np.ascontiguousarray(self.host)))
            self.where = "both"
        return self.dev

    def written(self):
        self.where = "dev"


def _field(name):
    def get(self):
        rp = self._replay
        if rp is not None and (name not in ("rho", "u") or rp.it != rp.it0):
            raise RuntimeError("during run(), app hooks may read lattice.rho and lattice.u "
                               "only at iterations where outputs are due")
        return self._fields[name].read_host()

    def put(self, value):
        self._fields[name].assign(value)

    return property(get, put)


def _ptr(t):
    return None if t is None else t.data_ptr()


class lattice:
    g = _field("g")
    g_eq = _field("g_eq")
    g_up = _field("g_up")
    rho = _field("rho")
    u = _field("u")

    # ------------------------------------------------------------ set-up --
    def __init__(self, app):
        # defaults, then the app's attributes (lattice.py:20-78)
        self.name = 'lattice'
        self.x_min = 0.0
        self.x_max = 1.0
        self.y_min = 0.0
        self.y_max = 1.0
        self.nx = 100
        self.ny = 100
        self.tau_lbm = 1.0
        self.dx = 1.0
        self.dt = 1.0
        self.Cx = self.dx
        self.Ct = self.dt
        self.Cr = 1.0
        self.Cn = self.Cx ** 2 / self.Ct
        self.Cu = self.Cx / self.Ct
        self.Cf = self.Cr * self.Cx ** 2 / self.Ct
        self.dpi = 100
        self.u_lbm = 0.03
        self.L_lbm = 100
        self.nu_lbm = 0.01
        self.Re_lbm = 100.0
        self.rho_lbm = 1.0
        self.IBB = False
        self.stop = 'it'
        self.t_max = 1.0
        self.it_max = 1000
        self.obs_cv_ct = 1.0e-1
        self.obs_cv_nb = 500
        self.precision = 'f64'
        for key in ("name", "x_min", "x_max", "y_min", "y_max", "nx", "ny", "tau_lbm", "dx",
                    "dt", "Cx", "Ct", "Cn", "Cu", "Cf", "dpi", "u_lbm", "L_lbm", "nu_lbm",
                    "Re_lbm", "rho_lbm", "IBB", "stop", "t_max", "it_max", "obs_cv_ct",
                    "obs_cv_nb", "precision"):
            if hasattr(app, key):
                setattr(self, key, getattr(app, key))
        if hasattr(app, "Cr"):
            self.Cr = app.Ct  # lattice.py:63 copies Ct into Cr (the value is unused)
        if self.precision not in _lib.STORE:
            raise ValueError("precision must be one of " + ", ".join(_lib.STORE))

        # output folders (lattice.py:80-91)
        time = datetime.now().strftime('%Y-%m-%d_%H_%M_%S')
        self.results_dir = getattr(app, "results_dir", './results/')
        self.output_dir = self.results_dir + str(time) + '/'
        self.png_dir = self.output_dir + './png/'
        os.makedirs(self.png_dir, exist_ok=True)

        self.set_default_lbm()
        self._recording = None      # list while app.set_bc is being recorded
        self._replay = None         # fused-run state seen by the app hooks
        self._dev_cache = {}

        if getattr(app, "verbose", True):
            print('##################')
            print('### LBM solver ###')
            print('##################')
            print('')
            print('# name       = ' + self.name)
            print('# u_lbm      = ' + '{:f}'.format(self.u_lbm))
            print('# L_lbm      = ' + '{:f}'.format(self.L_lbm))
            print('# nu_lbm     = ' + '{:f}'.format(self.nu_lbm))
            print('# Re_lbm     = ' + '{:f}'.format(self.Re_lbm))
            print('# tau_p_lbm  = ' + '{:f}'.format(self.tau_p_lbm))
            print('# tau_m_lbm  = ' + '{:f}'.format(self.tau_m_lbm))
            print('# dt         = ' + '{:f}'.format(self.dt))
            print('# dx         = ' + '{:f}'.format(self.dx))
            print('# nx         = ' + str(self.nx))
            print('# ny         = ' + str(self.ny))
            print('# IBB        = ' + str(self.IBB))
            print('')

    def set_default_lbm(self):
        """Constants and arrays (lattice.py:117-174)."""
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
        self.c = geometry.C.copy()
        self.w = np.array([4. / 9.] + [1. / 9.] * 4 + [1. / 36.] * 4)
        self.ns = geometry.NS.copy()
        shp = (self.q, self.nx, self.ny)
        self._fields = {
            "g": _DevField(np.zeros(shp)),
            "g_eq": _DevField(np.zeros(shp)),
            "g_up": _DevField(np.zeros(shp)),
            "rho": _DevField(np.ones((self.nx, self.ny))),
            "u": _DevField(np.zeros((2, self.nx, self.ny))),
        }
        self.u_left = np.zeros((2, self.ny))
        self.u_right = np.zeros((2, self.ny))
        self.u_top = np.zeros((2, self.nx))
        self.u_bot = np.zeros((2, self.nx))
        self.rho_right = np.zeros(self.ny)
        self.lattice = np.zeros((self.nx, self.ny))

    # ----------------------------------------------------------- helpers --
    def _need_gpu(self):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("lbm_amd runs its kernels on a ROCm GPU; none is visible")
        return _lib.load()

    def _side(self, arr):
        import torch
        return torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float64)).to("cuda")

    def _dev_int64(self, arr):
        import torch
        key = id(arr)
        hit = self._dev_cache.get(key)
        if hit is not None and hit[0] is arr:
            return hit[1]
        t = torch.from_numpy(np.ascontiguousarray(arr, dtype=np.int64)).to("cuda")
        self._dev_cache[key] = (arr, t)
        return t

    def _dev_f64(self, arr):
        import torch
        key = id(arr)
        hit = self._dev_cache.get(key)
        if hit is not None and hit[0] is arr:
            return hit[1]
        t = torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float64)).to("cuda")
        self._dev_cache[key] = (arr, t)
        return t

    def _check_replay(self):
        if self._replay is not None:
            raise RuntimeError("lattice kernels cannot be called from app hooks during run()")

    # ---------------------------------------------------------- hot path --
    def macro(self):
        """Moments rho, u of g (lattice.py:178-189) -> lbm_macro."""
        self._check_replay()
        L = self._need_gpu()
        f = self._fields
        g = f["g"].device()
        rho = f["rho"].device(load=False)
        u = f["u"].device(load=False)
        _lib.check(L.lbm_macro(self.nx, self.ny, _ptr(g), _ptr(rho), _ptr(u), _lib.stream_ptr()),
                   "lbm_macro")
        f["rho"].written()
        f["u"].written()

    def equilibrium(self):
        """g_eq from rho, u (nb.py:7-17) -> lbm_equilibrium."""
        self._check_replay()
        L = self._need_gpu()
        f = self._fields
        u, rho = f["u"].device(), f["rho"].device()
        geq = f["g_eq"].device(load=False)
        _lib.check(L.lbm_equilibrium(self.nx, self.ny, _ptr(u), _ptr(rho), _ptr(geq),
                                     _lib.stream_ptr()), "lbm_equilibrium")
        f["g_eq"].written()

    def collision_stream(self):
        """TRT collision and push streaming (nb.py:21-45) -> lbm_col_str."""
        self._check_replay()
        L = self._need_gpu()
        f = self._fields
        g, geq = f["g"].device(), f["g_eq"].device()
        gup = f["g_up"].device(load=False)
        _lib.check(L.lbm_col_str(self.nx, self.ny, _ptr(g), _ptr(geq), _ptr(gup),
                                 self.om_p_lbm, self.om_m_lbm, _lib.stream_ptr()), "lbm_col_str")
        f["g"].written()
        f["g_up"].written()

    def drag_lift(self, obs, R_ref, U_ref, L_ref):
        """Momentum-exchange drag and lift coefficients (nb.py:49-73)."""
        if self._replay is not None:
            fx, fy = self._replay.drag_sums(obs)
        else:
            import torch
            L = self._need_gpu()
            f = self._fields
            b = self._dev_int64(obs.boundary)
            out = torch.zeros(2, dtype=torch.float64, device="cuda")
            _lib.check(L.lbm_drag_lift(self.nx, self.ny, int(obs.boundary.shape[0]), _ptr(b),
                                       _ptr(f["g_up"].device()), _ptr(f["g"].device()),
                                       _ptr(out), _lib.stream_ptr()), "lbm_drag_lift")
            fx, fy = (float(v) for v in out.cpu().numpy())
        Cx = -2.0 * fx / (R_ref * L_ref * U_ref ** 2)
        Cy = -2.0 * fy / (R_ref * L_ref * U_ref ** 2)
        return Cx, Cy

    def drag_lift_chunk(self, obs, R_ref, U_ref, L_ref):
        """drag_lift of every iteration of the chunk run() is replaying (for an
        app's replay_chunk): arrays (Cx, Cy), each element computed as
        drag_lift computes it."""
        if self._replay is None or not hasattr(self._replay, "drag_sums_chunk"):
            raise RuntimeError("drag_lift_chunk is for replay_chunk hooks during run()")
        fx, fy = self._replay.drag_sums_chunk(obs)
        den = (R_ref * L_ref * U_ref ** 2)
        return -2.0 * fx / den, -2.0 * fy / den

    def bounce_back_obstacle(self, obstacle):
        """Plain or interpolated bounce-back on an obstacle's links (nb.py:77-117)."""
        if self._recording is not None:
            self._recording.append(("obstacle", obstacle))
            return
        self._check_replay()
        L = self._need_gpu()
        bnd = np.asarray(obstacle.boundary)
        nb = int(bnd.shape[0])
        if nb == 0:
            return
        for ax, n in ((0, self.nx), (1, self.ny)):
            if bnd[:, ax].min() < -n or bnd[:, ax].max() >= n:
                raise IndexError("obstacle boundary index out of range")
        f = self._fields
        b = self._dev_int64(obstacle.boundary)
        ibb = self._dev_f64(obstacle.ibb) if self.IBB else None
        _lib.check(L.lbm_bounce_back_obstacle(self.nx, self.ny, int(bool(self.IBB)), nb, _ptr(b),
                                              _ptr(ibb), _ptr(f["g_up"].device()),
                                              _ptr(f["g"].device()), _lib.stream_ptr()),
                   "lbm_bounce_back_obstacle")
        f["g"].written()

    def _wall(self, kind, fn, side, extra=None):
        if self._recording is not None:
            self._recording.append((kind, None))
            return
        n = self.nx if kind in ("top", "bottom") else self.ny
        if np.shape(side) != (2, n):
            raise ValueError(f"zou_he {kind}: wall velocity has shape {np.shape(side)}, "
                             f"expected {(2, n)}")
        if extra is not None and np.shape(extra) != (self.ny,):
            raise ValueError(f"zou_he {kind}: rho_right has shape {np.shape(extra)}, "
                             f"expected {(self.ny,)}")
        self._check_replay()
        L = self._need_gpu()
        f = self._fields
        u, rho, g = f["u"].device(), f["rho"].device(), f["g"].device()
        s = self._side(side)
        st = _lib.stream_ptr()
        if extra is None:
            rc = getattr(L, fn)(self.nx, self.ny, _ptr(u), _ptr(s), _ptr(rho), _ptr(g), st)
        else:
            e = self._side(extra)
            rc = getattr(L, fn)(self.nx, self.ny, _ptr(u), _ptr(e), _ptr(s), _ptr(rho), _ptr(g), st)
        _lib.check(rc, fn)
        f["u"].written()
        f["rho"].written()
        f["g"].written()

    def zou_he_left_wall_velocity(self):
        """nb.py:121-143"""
        self._wall("left", "lbm_zou_he_left_wall_velocity", self.u_left)

    def zou_he_right_wall_velocity(self):
        """nb.py:147-169"""
        self._wall("right", "lbm_zou_he_right_wall_velocity", self.u_right)

    def zou_he_right_wall_pressure(self):
        """nb.py:173-195"""
        self._wall("right_pressure", "lbm_zou_he_right_wall_pressure", self.u_right,
                   extra=self.rho_right)

    def zou_he_top_wall_velocity(self):
        """nb.py:199-221"""
        self._wall("top", "lbm_zou_he_top_wall_velocity", self.u_top)

    def zou_he_bottom_wall_velocity(self):
        """nb.py:225-247"""
        self._wall("bottom", "lbm_zou_he_bottom_wall_velocity", self.u_bot)

    def _corner(self, kind, fn):
        if self._recording is not None:
            self._recording.append((kind, None))
            return
        self._check_replay()
        L = self._need_gpu()
        f = self._fields
        u, rho, g = f["u"].device(), f["rho"].device(), f["g"].device()
        _lib.check(getattr(L, fn)(self.nx, self.ny, _ptr(u), _ptr(rho), _ptr(g),
                                  _lib.stream_ptr()), fn)
        f["u"].written()
        f["rho"].written()
        f["g"].written()

    def zou_he_bottom_left_corner(self):
        """nb.py:251-271"""
        self._corner("bottom_left", "lbm_zou_he_bottom_left_corner")

    def zou_he_top_left_corner(self):
        """nb.py:275-296"""
        self._corner("top_left", "lbm_zou_he_top_left_corner")

    def zou_he_top_right_corner(self):
        """nb.py:300-320"""
        self._corner("top_right", "lbm_zou_he_top_right_corner")

    def zou_he_bottom_right_corner(self):
        """nb.py:324-344"""
        self._corner("bottom_right", "lbm_zou_he_bottom_right_corner")

    # ------------------------------------------------------------- setup --
    def add_obstacle(self, obstacle):
        """Rasterise, build links and IBB distances (lattice.py:290-375)."""
        return geometry.add_obstacle(self, obstacle)

    def get_coords(self, i, j):
        """Coordinates of node (i, j) (lattice.py:379-387)."""
        dx = (self.x_max - self.x_min) / (self.nx - 1)
        dy = (self.y_max - self.y_min) / (self.ny - 1)
        return [self.x_min + i * dx, self.y_min + j * dy]

    def is_inside(self, poly, pt):
        """Even-odd inside test of one point (lattice.py:391-414)."""
        return bool(geometry.inside_polygon(poly, np.array([pt[0]]), np.array([pt[1]]))[0])

    def generate_image(self, obstacles):
        """Obstacle map with link nodes marked (lattice.py:418-436)."""
        import matplotlib
        matplotlib.use("Agg")
        import matplotlib.pyplot as plt
        lat = self.lattice.astype(float)
        for obs in obstacles:
            lat[obs.boundary[:, 0], obs.boundary[:, 1]] = -1.0
        plt.imsave(self.output_dir + self.name + '.png', np.rot90(lat), vmin=-1.0, vmax=1.0)

    # ----------------------------------------------------- extension API --
    def device_fields(self):
        """The float64 device tensors (g, g_eq, g_up, rho, u), synchronised."""
        self._need_gpu()
        return {k: f.device() for k, f in self._fields.items()}
