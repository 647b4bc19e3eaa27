"""The time loop (lbm/src/core/run.py:12-61) over the fused HIP time step.

``run(lattice, app)`` keeps the reference's contract: the app's hooks are
called in the reference order for every iteration, and on return the lattice
holds exactly what the reference leaves (g, g_up and g_eq of the last
iteration, rho/u = moments of the last iteration with the Zou-He values on
the perimeter).  Underneath, iterations run in chunks on the GPU:

1. ``app.set_bc`` is recorded once into a boundary-condition plan (obstacle
   links, walls, corners); the per-call kernels are not launched.
2. For a chunk of iterations [it0, it1): ``app.set_inlets`` is called for
   each of them and the wall arrays are captured into an inlet table; the
   iterations then run as fused launches (one per iteration, replayed from a
   HIP graph), with the drag/lift terms of every iteration kept on the device.
3. The hooks of those iterations are then replayed on the host in order:
   printings, outputs (which see rho/u of it0, where chunks start), observables
   (``lattice.drag_lift`` returns the value of that iteration), check_stop.
4. If check_stop ends the run inside the chunk, the state is restored from
   the chunk-start checkpoint and re-advanced to the stopping iteration.
5. Failure detection: after every chunk the FP64 density sum of the state
   (lbm_mass) must be finite, else InstabilityError is raised; an app
   attribute ``max_mass_drift`` also bounds its relative drift.  The sums
   are kept in ``lattice.mass_history``.
Chunks start at multiples of ``app.output_freq``; an app without it runs one
iteration per chunk.  A boundary plan the fused kernel cannot express runs the
reference's call-by-call loop on the per-call kernels instead.

Apps whose wall inputs are amplitude(it) x fixed profile (the reference's
cavity and channel apps) declare it with ``separable_inlets(lattice)``: the
wall arrays then come from device tables (one amplitude row per iteration,
lbm_fused_set_amplitudes) instead of a ``set_inlets`` call and a snapshot per
iteration; ``set_inlets`` still runs once per chunk, and its arrays must equal
amplitude x profile bit for bit (else the run stops with an error).  Such runs
are pipelined: chunk k+1 runs on the GPU while the hooks of chunk k replay on
the host (a stop inside chunk k discards it and re-advances chunk k from its
checkpoint).  An app may also replay a whole chunk at once with
``replay_chunk(lattice, it0, it1, at)`` (vectorised observables; ``at(it)``
marks the iteration whose hooks run), returning the stopping iteration or None.
"""
import math
import time

import numpy as np

from lbm_amd import _lib
from lbm_amd.domain import FusedDomain, q16_range

_SIDES = ("u_bot", "u_left", "u_top", "u_right", "rho_right")
_WALLS = {"bottom": 0, "left": 1, "top": 2, "right": 3, "right_pressure": 3}
_CORNERS = ("bottom_left", "top_left", "top_right", "bottom_right")


class PlanError(ValueError):
    """The recorded set_bc plan has no fused equivalent."""


class InstabilityError(FloatingPointError):
    """The populations stopped being finite (or the density drifted beyond
    the app's ``max_mass_drift``) during a fused run."""


def record_plan(lattice, app):
    lattice._recording = []
    try:
        app.set_bc(lattice)
        return lattice._recording
    finally:
        lattice._recording = None


def compile_plan(plan):
    """Map a recorded set_bc sequence onto the fused kernel's plan."""
    bc = [0, 0, 0, 0]
    corners = set()
    obstacles = []
    seen_wall = False
    obs_before = obs_after = False
    for kind, arg in plan:
        if kind == "obstacle":
            if any(o is arg for o in obstacles):
                raise PlanError("an obstacle appears twice in set_bc")
            obstacles.append(arg)
            if corners:
                raise PlanError("obstacle after the corners")
            if seen_wall:
                obs_after = True
            else:
                obs_before = True
        elif kind in _WALLS:
            if corners:
                raise PlanError("wall after a corner")
            side = _WALLS[kind]
            if bc[side]:
                raise PlanError("a wall appears twice in set_bc")
            bc[side] = _lib.BC_PRESSURE if kind == "right_pressure" else _lib.BC_VELOCITY
            seen_wall = True
        elif kind in _CORNERS:
            if kind in corners:
                raise PlanError("a corner appears twice in set_bc")
            corners.add(kind)
        else:
            raise PlanError("unknown boundary call " + kind)
    if obs_before and obs_after:
        raise PlanError("obstacles both before and after the walls")
    if not all(bc) or len(corners) != 4:
        raise PlanError("the fused plan needs all four walls and all four corners")
    return bc, obstacles, (not obs_after)


class _InletLog:
    """Deduplicated snapshots of the five wall arrays, one row per iteration."""

    def __init__(self, lattice):
        self.lat = lattice
        self.last = None      # tuple of arrays of the previous iteration

    def snapshot(self):
        cur = []
        for i, name in enumerate(_SIDES):
            a = getattr(self.lat, name)
            prev = None if self.last is None else self.last[i]
            if prev is not None and prev.shape == a.shape and np.array_equal(prev, a):
                cur.append(prev)
            else:
                cur.append(np.array(a, dtype=np.float64, copy=True))
        self.last = tuple(cur)
        return self.last


def app_q16_range(app, lattice):
    """Range of f_q - w_q for the Q16 storage form: the app's ``q16_range``
    ([9][2]) if it sets one, else q16_range(u_bound) with the app's
    ``q16_u_bound`` or 1.5 x its velocity scale u_lbm (the lid or inlet peak
    speed of the reference apps).  None for the other forms."""
    if lattice.precision != "q16":
        return None
    r = getattr(app, "q16_range", None)
    if r is not None:
        return r
    u = getattr(app, "q16_u_bound", None)
    if u is None:
        u = 1.5 * float(getattr(app, "u_lbm", lattice.u_lbm))
    return q16_range(u)


def _separable(app, lattice):
    """(profiles, amplitudes) of an app that declares separable wall inputs,
    else None: profiles = five arrays (u_bot, u_left, u_top, u_right,
    rho_right), amplitudes(it) = five factors."""
    fn = getattr(app, "separable_inlets", None)
    if fn is None:
        return None
    # only when set_inlets is the one the declaration describes (a subclass
    # overriding set_inlets alone keeps the per-iteration path)
    owner = next(c for c in type(app).__mro__ if "separable_inlets" in c.__dict__)
    if type(app).set_inlets is not owner.__dict__.get("set_inlets"):
        return None
    profiles, amp = fn(lattice)
    shapes = ((2, lattice.nx), (2, lattice.ny), (2, lattice.nx), (2, lattice.ny), (lattice.ny,))
    prof = []
    for p, shp in zip(profiles, shapes):
        a = np.zeros(shp) if p is None else np.ascontiguousarray(p, dtype=np.float64)
        if a.shape != shp:
            raise ValueError(f"separable_inlets: profile of shape {a.shape}, expected {shp}")
        prof.append(a)
    return prof, amp


class FusedRun:
    """Drives ``app`` on ``lattice`` with the fused time step."""

    def __init__(self, lattice, app, chunk=1000, cells_per_thread=0):
        import torch
        self.lat = lattice
        self.app = app
        self.chunk = int(chunk)
        lattice._need_gpu()
        bc, obstacles, obs_first = compile_plan(record_plan(lattice, app))
        self.obstacles = obstacles
        self.dom = FusedDomain(lattice.nx, lattice.ny, lattice.om_p_lbm, lattice.om_m_lbm, bc,
                               [(o.boundary, getattr(o, "ibb", None)) for o in obstacles],
                               ibb=bool(lattice.IBB), obstacles_first=obs_first,
                               storage=lattice.precision, ring=self.chunk + 2,
                               cells_per_thread=cells_per_thread,
                               q16_range=app_q16_range(app, lattice))
        self.log = _InletLog(lattice)
        self.sep = _separable(app, lattice)
        if self.sep is not None:
            self.dom.set_profiles(self.sep[0])
        self.mass_history = []       # (iteration, total density in FP64) per chunk
        self.max_drift = getattr(app, "max_mass_drift", None)
        self.prev_row = None
        self.drag = None
        self.drag_it0 = 0
        n = (lattice.nx, lattice.ny)
        self.rho_b = torch.empty(n, dtype=torch.float64, device="cuda")
        self.u_b = torch.empty((2,) + n, dtype=torch.float64, device="cuda")
        # pipelined runs: moments of the first iteration of the chunk ahead
        self.rho_s = torch.empty(n, dtype=torch.float64, device="cuda") if self.sep else None
        self.u_s = torch.empty((2,) + n, dtype=torch.float64, device="cuda") if self.sep else None

    def close(self):
        self.dom.close()

    def _upload(self, it0, rows):
        """Inlet rows for iterations it0-1 (repeated from the last chunk) .. it0+len-1."""
        all_rows = ([self.prev_row] if it0 > 0 else []) + rows
        self.dom.set_inlet_rows(it0 - 1 if it0 > 0 else 0, all_rows)
        self.prev_row = rows[-1]

    def _upload_sep(self, it0, it1):
        """Amplitude rows of iterations it0-1 .. it1-1 (separable inlets)."""
        amp = self.sep[1]
        r0 = max(it0 - 1, 0)
        rows = np.array([amp(it) for it in range(r0, it1)], dtype=np.float64).reshape(-1, 5)
        self.dom.set_amplitudes(r0, rows)

    def _check_separable(self, it):
        """The app's set_inlets(it) must give amplitude x profile exactly."""
        self.app.set_inlets(self.lat, it)
        prof, amp = self.sep
        a = amp(it)
        for k, name in enumerate(_SIDES):
            ref = np.asarray(getattr(self.lat, name), dtype=np.float64)
            if not np.array_equal(ref, a[k] * prof[k]):
                raise RuntimeError(f"separable_inlets of {type(self.app).__name__} disagrees "
                                   f"with its set_inlets at iteration {it} ({name})")

    def _advance(self, it0, it1, macro_last, first=None):
        """Run iterations it0..it1-1: macro of it0 into lattice rho/u (or into
        ``first`` = (rho, u)), of it1-1 into the chunk-end buffers (or lattice
        rho/u when macro_last)."""
        dom = self.dom
        f = self.lat._fields
        if first is None:
            rho, u = f["rho"].device(load=False), f["u"].device(load=False)
        else:
            rho, u = first
        if it0 == 0:
            dom.init(f["g"].device(), rho, u)
        else:
            dom.step(it0, rho, u)
        if first is None:
            f["rho"].written()
            f["u"].written()
        if it1 - it0 >= 3:
            dom.steps(it0 + 1, it1 - it0 - 2)
        if it1 - it0 >= 2:
            rb, ub = (rho, u) if macro_last else (self.rho_b, self.u_b)
            dom.step(it1 - 1, rb, ub)

    def _finish(self, it_last):
        """Epilogue of the last iteration: g, g_up, g_eq, rho/u as the reference leaves them."""
        f = self.lat._fields
        rho, u = f["rho"].device(), f["u"].device()
        geq = f["g_eq"].device(load=False)
        _lib.check(self.dom.L.lbm_equilibrium(self.lat.nx, self.lat.ny, u.data_ptr(),
                                              rho.data_ptr(), geq.data_ptr(), _lib.stream_ptr()),
                   "lbm_equilibrium")
        self.dom.epilogue(it_last, f["g"].device(load=False), f["g_up"].device(load=False),
                          rho, u)
        for k in ("g_eq", "g", "g_up", "rho", "u"):
            f[k].written()

    def drag_sums(self, obs):
        it = self.lat._replay.it
        for k, o in enumerate(self.obstacles):
            if o is obs:
                row = self.drag[it - self.drag_it0, k]
                return float(row[0]), float(row[1])
        raise KeyError("drag_lift on an obstacle that set_bc does not bounce back")

    # -- loop
    def _chunk_end(self, it0):
        app = self.app
        it1 = it0 + self.chunk
        freq = getattr(app, "output_freq", None)
        if isinstance(freq, (int, np.integer)) and freq > 0:
            it1 = min(it1, (it0 // freq + 1) * freq)
        else:
            it1 = it0 + 1
        if getattr(app, "stop", None) == "it" and hasattr(app, "it_max"):
            it1 = min(it1, int(app.it_max) + 1)
        return max(it1, it0 + 1)

    def go(self):
        if self.sep is not None:
            return self._go_pipelined()
        app, lat, dom = self.app, self.lat, self.dom
        it0 = 0
        while True:
            it1 = self._chunk_end(it0)
            rows = []
            for it in range(it0, it1):
                app.set_inlets(lat, it)
                rows.append(self.log.snapshot())
            self._upload(it0, rows)
            if it0 > 0:
                dom.save(it0 - 1)
            self._advance(it0, it1, macro_last=False)
            if self.obstacles:
                dom.drag_only(it1 - 1)
                self.drag, self.drag_it0 = dom.drag_read(it0, it1 - it0), it0
            self._check_state(it1 - 1)
            stop_it = self._replay(it0, it1)
            if stop_it is None:
                it0 = it1
                continue
            if stop_it < it1 - 1:
                if it0 > 0:
                    dom.restore(it0 - 1)
                self._advance(it0, stop_it + 1, macro_last=True)
            elif it1 - it0 >= 2:
                f = lat._fields
                f["rho"].device(load=False).copy_(self.rho_b)
                f["u"].device(load=False).copy_(self.u_b)
            self._finish(stop_it)
            return stop_it

    def _last_chunk(self, it1):
        """True when check_stop must end the run inside [.., it1)."""
        app = self.app
        return getattr(app, "stop", None) == "it" and hasattr(app, "it_max") and \
            it1 >= int(app.it_max) + 1

    def _go_pipelined(self):
        import torch
        lat, dom = self.lat, self.dom
        it0, it1 = 0, self._chunk_end(0)
        self._check_separable(0)
        self._upload_sep(it0, it1)
        self._advance(it0, it1, macro_last=False)
        while True:
            if self.obstacles:
                dom.drag_only(it1 - 1)
                self.drag, self.drag_it0 = dom.drag_read(it0, it1 - it0), it0
            self._check_state(it1 - 1)
            nxt = None
            if not self._last_chunk(it1):
                # the next chunk runs on the GPU while this one's hooks replay
                n1 = self._chunk_end(it1)
                self._check_separable(it1)
                self._upload_sep(it1, n1)
                dom.save(it1 - 1)
                self._advance(it1, n1, macro_last=False, first=(self.rho_s, self.u_s))
                nxt = (it1, n1)
            stop_it = self._replay(it0, it1)
            if stop_it is None:
                if nxt is None:
                    raise RuntimeError("check_stop did not end the run at it_max")
                f = lat._fields
                f["rho"].device(load=False).copy_(self.rho_s)
                f["u"].device(load=False).copy_(self.u_s)
                f["rho"].written()
                f["u"].written()
                it0, it1 = nxt
                continue
            # stop inside this chunk: drop the chunk ahead, re-advance this one
            torch.cuda.synchronize()
            if it0 > 0:
                dom.restore(it0 - 1)
            self._upload_sep(it0, stop_it + 1)
            self._advance(it0, stop_it + 1, macro_last=True)
            self._finish(stop_it)
            # the wall arrays as the reference leaves them
            self.app.set_inlets(lat, stop_it)
            return stop_it

    def _check_state(self, it):
        """Failure detection once per chunk: the FP64 density sum of the
        post-collision state of iteration ``it`` (lbm_mass) must be finite --
        a NaN or an infinity anywhere in the populations makes it
        non-finite -- and, if the app sets ``max_mass_drift``, stay within
        that relative distance of the first chunk's value."""
        m = self.dom.mass(it)
        self.mass_history.append((it, m))
        if not math.isfinite(m):
            raise InstabilityError(f"non-finite populations after iteration {it} "
                                   f"(density sum {m})")
        if self.max_drift is not None:
            m0 = self.mass_history[0][1]
            drift = abs(m - m0) / abs(m0)
            if drift > self.max_drift:
                raise InstabilityError(f"density drift {drift:.3e} after iteration {it} "
                                       f"exceeds max_mass_drift = {self.max_drift}")

    def _replay(self, it0, it1):
        app, lat = self.app, self.lat
        state = _Replay(self, it0)
        state.it1 = it1
        lat._replay = state
        try:
            chunk_fn = getattr(app, "replay_chunk", None)
            if chunk_fn is not None:
                def at(it):
                    state.it = it
                stop = chunk_fn(lat, it0, it1, at)
                if stop is not NotImplemented:
                    return stop
            for it in range(it0, it1):
                state.it = it
                app.printings(it)
                app.outputs(lat, it)
                app.observables(lat, it)
                if not app.check_stop(it):
                    return it
            return None
        finally:
            lat._replay = None


class _Replay:
    def __init__(self, run, it0):
        self.run = run
        self.it0 = it0
        self.it = it0

    def drag_sums(self, obs):
        return self.run.drag_sums(obs)

    def drag_sums_chunk(self, obs):
        """(fx, fy) arrays of the iterations it0 .. it1-1 of the chunk."""
        run = self.run
        for k, o in enumerate(run.obstacles):
            if o is obs:
                rows = run.drag[self.it0 - run.drag_it0:self.it1 - run.drag_it0, k]
                return rows[:, 0].copy(), rows[:, 1].copy()
        raise KeyError("drag_lift on an obstacle that set_bc does not bounce back")


def _run_per_call(lattice, app):
    """The reference loop on the per-call kernels (run.py:24-54)."""
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
        compute = app.check_stop(it)
        it += 1
    return it - 1


def run(lattice, app, chunk=1000, fused=True):
    """Run ``app`` on ``lattice`` (run.py:12-61); returns the last iteration."""
    app.initialize(lattice)
    start = time.time()
    print('### Solving')
    engine = None
    if fused:
        try:
            engine = FusedRun(lattice, app, chunk=chunk)
        except (PlanError, _lib.Unsupported) as e:
            print('# fused time step unavailable (' + str(e) + '), using per-call kernels')
            engine = None
    if engine is not None:
        try:
            last = engine.go()
        finally:
            lattice.mass_history = engine.mass_history
            engine.close()
    else:
        last = _run_per_call(lattice, app)
    lattice.loop_time = time.time() - start
    print("# Loop time = {:f}".format(lattice.loop_time))
    app.finalize(lattice)
    return last
