"""FusedDomain: one lattice's fused time step on the GPU (lbm_fused_* of
include/lbm_d2q9.h) with its boundary plan, obstacle links and inlet table.

Used by ``lbm_amd.run`` (the reference's time loop) and directly by callers
that own their loop (benchmarks, batch drivers).  Device pointers passed in
are torch float64 CUDA tensors with the reference layout.
"""
import ctypes

import numpy as np

from lbm_amd import _lib


_C = np.array([[0, 0], [1, 0], [-1, 0], [0, 1], [0, -1], [1, 1], [-1, -1], [-1, 1], [1, -1]])
_W = np.array([4. / 9.] + [1. / 9.] * 4 + [1. / 36.] * 4)


def q16_range(u_bound, rho_bound=0.05, margin=1.25):
    """Per-direction range [lo, hi] of h_q = f_q - w_q for the Q16 storage form:
    the equilibrium deviation w_q (drho + rho (3 c.u + 4.5 (c.u)^2 - 1.5 u.u))
    over |u| <= u_bound and |rho - 1| <= rho_bound, widened about its centre by
    ``margin`` for the non-equilibrium part.  Returns float64 [9][2]."""
    U, D = float(u_bound), float(rho_bound)
    out = np.zeros((9, 2))
    for q in range(9):
        a = float(np.hypot(*_C[q])) * U            # largest |c_q . u|
        top = 3.0 * a + 4.5 * a * a                # largest 3 cu + 4.5 cu^2 (u.u >= 0)
        bot = -3.0 * a - 1.5 * U * U               # smallest 3 cu + 4.5 cu^2 - 1.5 u.u
        lo = -D + (1.0 + D) * min(bot, 0.0)
        hi = D + (1.0 + D) * max(top, 0.0)
        c, r = 0.5 * (lo + hi), 0.5 * (hi - lo) * float(margin)
        out[q] = (_W[q] * (c - r), _W[q] * (c + r))
    return out


class FusedDomain:
    def __init__(self, nx, ny, om_p, om_m, bc, obstacles_links=(), ibb=False,
                 obstacles_first=True, storage="f64", ring=1002, cells_per_thread=0,
                 vector_width=0, q16_range=None):
        """bc: four LBM_BC_* codes (bottom, left, top, right).
        obstacles_links: sequence of (boundary int[nb,3], ibb float[nb] or None).
        q16_range: [9][2] range of f_q - w_q for storage "q16" (see q16_range())."""
        self.L = _lib.load()
        self.nx, self.ny = int(nx), int(ny)
        self.storage = storage
        self.ring = int(ring)
        links, dists, modes, offs = [], [], [], [0]
        for bnd, dist in obstacles_links:
            b = np.asarray(bnd, dtype=np.int64).reshape(-1, 3)
            links.append(b)
            dists.append(np.zeros(len(b)) if (not ibb or dist is None)
                         else np.asarray(dist, dtype=np.float64).reshape(-1))
            modes.append(np.full(len(b), 1 if ibb else 0, dtype=np.uint8))
            offs.append(offs[-1] + len(b))
        self.n_obs = len(offs) - 1
        self.links = np.ascontiguousarray(np.concatenate(links) if links
                                          else np.zeros((0, 3), np.int64))
        self.dists = np.ascontiguousarray(np.concatenate(dists) if dists else np.zeros(0))
        self.modes = np.ascontiguousarray(np.concatenate(modes) if modes
                                          else np.zeros(0, np.uint8))
        self.obs_off = np.asarray(offs, dtype=np.int32)
        d = _lib.FusedDesc()
        d.nx, d.ny = self.nx, self.ny
        d.storage = _lib.STORE[storage]
        d.om_p, d.om_m = float(om_p), float(om_m)
        for i in range(4):
            d.bc[i] = int(bc[i])
        d.corners = 1
        d.obstacles_first = int(bool(obstacles_first))
        d.n_links = len(self.links)
        d.links_host = self.links.ctypes.data_as(ctypes.POINTER(_lib.c_i64))
        d.ibb_host = self.dists.ctypes.data_as(_lib.c_dp)
        d.link_ibb_host = self.modes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        d.n_obs = self.n_obs
        d.obs_offsets_host = self.obs_off.ctypes.data_as(ctypes.POINTER(_lib.c_i32))
        d.ring = self.ring
        d.cells_per_thread = int(cells_per_thread)
        d.vector_width = int(vector_width)
        if q16_range is not None:
            self.q16 = np.ascontiguousarray(q16_range, dtype=np.float64).reshape(9, 2)
            d.q16_range_host = self.q16.ctypes.data_as(_lib.c_dp)
        h = _lib.c_vp()
        _lib.check(self.L.lbm_fused_create(ctypes.byref(d), ctypes.byref(h)), "lbm_fused_create")
        self.h = h

    # ------------------------------------------------------------ control --
    def close(self):
        if getattr(self, "h", None):
            self.L.lbm_fused_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _p(t):
        return None if t is None else t.data_ptr()

    def set_inlets(self, row0, offsets, pool):
        offs = np.ascontiguousarray(offsets, dtype=np.int32)
        pool = np.ascontiguousarray(pool, dtype=np.float64)
        _lib.check(self.L.lbm_fused_set_inlets(self.h, int(row0), len(offs),
                                               offs.ctypes.data_as(_lib.c_vp), pool.size,
                                               pool.ctypes.data_as(_lib.c_vp), _lib.stream_ptr()),
                   "lbm_fused_set_inlets")

    def set_inlet_rows(self, row0, rows):
        """Rows row0.. of the inlet table from snapshots of the five wall
        arrays, each checked against this lattice's side shapes."""
        pool, offs = pack_inlets(rows, self.nx, self.ny)
        self.set_inlets(row0, offs, pool)

    def set_profiles(self, profiles):
        """Separable wall inputs: side s at iteration it = amp[it][s] * profiles[s]
        (five arrays: u_bot [2][nx], u_left [2][ny], u_top [2][nx], u_right
        [2][ny], rho_right [ny]; None = zeros)."""
        keep = [None if p is None else np.ascontiguousarray(p, dtype=np.float64) for p in profiles]
        arr = (_lib.c_dp * 5)(*[(_lib.c_dp() if p is None else p.ctypes.data_as(_lib.c_dp))
                                for p in keep])
        _lib.check(self.L.lbm_fused_set_profiles(self.h, arr, _lib.stream_ptr()),
                   "lbm_fused_set_profiles")

    def set_amplitudes(self, row0, amp):
        """Rows row0 .. of the amplitude table: amp [n][5]."""
        a = np.ascontiguousarray(amp, dtype=np.float64).reshape(-1, 5)
        _lib.check(self.L.lbm_fused_set_amplitudes(self.h, int(row0), len(a),
                                                   a.ctypes.data_as(_lib.c_vp),
                                                   _lib.stream_ptr()),
                   "lbm_fused_set_amplitudes")

    def init(self, g, rho=None, u=None):
        _lib.check(self.L.lbm_init(self.h, self._p(g), self._p(rho), self._p(u),
                                   _lib.stream_ptr()), "lbm_init")

    def step(self, it, rho=None, u=None):
        _lib.check(self.L.lbm_step(self.h, int(it), self._p(rho), self._p(u), _lib.stream_ptr()),
                   "lbm_step")

    def steps(self, it0, n):
        _lib.check(self.L.lbm_steps(self.h, int(it0), int(n), _lib.stream_ptr()), "lbm_steps")

    def prepare(self, it0, n):
        """Capture and upload the graph steps(it0, n) will replay, without running it."""
        _lib.check(self.L.lbm_steps_prepare(self.h, int(it0), int(n), _lib.stream_ptr()),
                   "lbm_steps_prepare")

    def epilogue(self, it, g, g_up, rho=None, u=None):
        _lib.check(self.L.lbm_epilogue(self.h, int(it), self._p(g), self._p(g_up), self._p(rho),
                                       self._p(u), _lib.stream_ptr()), "lbm_epilogue")

    def drag_only(self, it):
        _lib.check(self.L.lbm_drag_only(self.h, int(it), _lib.stream_ptr()), "lbm_drag_only")

    def drag_read(self, it0, n):
        """(fx, fy) sums of iterations it0..it0+n-1: array [n][n_obs][2]."""
        out = np.zeros((int(n), max(self.n_obs, 1), 2))
        if self.n_obs:
            _lib.check(self.L.lbm_drag_read(self.h, int(it0), int(n),
                                            out.ctypes.data_as(_lib.c_vp), _lib.stream_ptr()),
                       "lbm_drag_read")
        return out[:, :self.n_obs]

    def save(self, it):
        _lib.check(self.L.lbm_fused_save(self.h, int(it), _lib.stream_ptr()), "lbm_fused_save")

    def restore(self, it):
        _lib.check(self.L.lbm_fused_restore(self.h, int(it), _lib.stream_ptr()),
                   "lbm_fused_restore")

    def mass(self, it):
        out = np.zeros(1)
        _lib.check(self.L.lbm_mass(self.h, int(it), out.ctypes.data_as(_lib.c_vp),
                                   _lib.stream_ptr()), "lbm_mass")
        return float(out[0])

    def device_bytes(self):
        return int(self.L.lbm_fused_device_bytes(self.h))


SIDE_NAMES = ("u_bot", "u_left", "u_top", "u_right", "rho_right")


def side_shapes(nx, ny):
    """Shapes of the five wall arrays (lattice.py:164-174)."""
    return ((2, nx), (2, ny), (2, nx), (2, ny), (ny,))


def pack_inlets(rows, nx=None, ny=None):
    """Pool (float64) and offsets [nrows][5] for snapshots of the five wall
    arrays (u_bot, u_left, u_top, u_right, rho_right); identical arrays
    (same object) are stored once.  With nx and ny given, every array must
    have its side's shape (ValueError otherwise)."""
    want = side_shapes(nx, ny) if nx is not None and ny is not None else None
    pool, offs, where = [np.zeros(1)], [], {}
    n = 1          # offset 0 is a zero pad
    for r, snap in enumerate(rows):
        if len(snap) != 5:
            raise ValueError(f"inlet row {r}: {len(snap)} arrays, expected 5")
        o = []
        for sd, a in enumerate(snap):
            if want is not None and np.shape(a) != want[sd]:
                raise ValueError(f"inlet row {r}: {SIDE_NAMES[sd]} has shape {np.shape(a)}, "
                                 f"expected {want[sd]}")
            k = id(a)
            if k not in where:
                where[k] = n
                flat = np.ascontiguousarray(a, dtype=np.float64).ravel()
                pool.append(flat)
                n += flat.size
            o.append(where[k])
        offs.append(o)
    return np.ascontiguousarray(np.concatenate(pool)), np.asarray(offs, dtype=np.int32)
