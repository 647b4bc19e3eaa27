"""Schaefer-Turek cylinder benchmark (lbm/src/app/turek.py:12-173)."""
import math

from lbm_amd.app._channel import channel_app
from lbm_amd.buff import buff
from lbm_amd.obstacle import obstacle


class turek(channel_app):
    def __init__(self):
        self.name = 'turek'
        self.Re_lbm = 100.0
        self.L_lbm = 151
        self.u_lbm = 0.05
        self.rho_lbm = 1.0
        self.t_max = 0.02
        self.x_min = -0.2
        self.x_max = 2.0
        self.y_min = -0.2
        self.y_max = 0.21
        self.IBB = True
        self.stop = 'obs'
        self.obs_cv_ct = 1.0e-3
        self.obs_cv_nb = 1000
        self.output_freq = 500
        self.output_it = 0
        self.dpi = 200
        self.compute_lbm_parameters()
        cylinder = obstacle('turek', 200, 2, 'cylinder', self.r_cyl, [0.0, 0.0])
        self.obstacles = [cylinder]

    def compute_lbm_parameters(self):
        self.Cs = 1.0 / math.sqrt(3.0)
        self.ny = self.L_lbm
        self.u_avg = 2.0 * self.u_lbm / 3.0
        self.r_cyl = 0.1
        self.D_lbm = math.floor(self.ny * self.r_cyl / (self.y_max - self.y_min))
        self.nu_lbm = self.u_avg * self.D_lbm / self.Re_lbm
        self.tau_lbm = 0.5 + self.nu_lbm / (self.Cs ** 2)
        self.dt = self.Re_lbm * self.nu_lbm / self.D_lbm ** 2
        self.dx = (self.y_max - self.y_min) / self.ny
        self.dy = self.dx
        self.nx = math.floor(self.ny * (self.x_max - self.x_min) / (self.y_max - self.y_min))
        self.it_max = math.floor(self.t_max / self.dt)
        self.sigma = math.floor(10 * self.nx)

    def initialize(self, lattice):
        self.init_obstacle_fields(lattice)
        self.drag_buff = buff('drag', lattice.dt, lattice.obs_cv_ct, lattice.obs_cv_nb,
                              lattice.output_dir)
        self.lift_buff = buff('lift', lattice.dt, lattice.obs_cv_ct, lattice.obs_cv_nb,
                              lattice.output_dir)
        lattice.equilibrium()
        lattice.g = lattice.g_eq.copy()
        self._dl_file = open(lattice.output_dir + 'drag_lift', 'a')

    def observables(self, lattice, it):
        drag, lift = lattice.drag_lift(self.obstacles[0], self.rho_lbm, self.u_avg, self.D_lbm)
        self.add_buff(drag, lift, lattice, it)

    def add_buff(self, Cx, Cy, lattice, it):
        """History, moving averages and the drag_lift file (turek.py:147-162)."""
        self.drag_buff.add(Cx)
        self.lift_buff.add(Cy)
        avg_Cx, dcx = self.drag_buff.mv_avg()
        avg_Cy, dcy = self.lift_buff.mv_avg()
        self._dl_file.write('{} {} {} {} {} {} {}\n'.format(it * self.dt, Cx, Cy, avg_Cx,
                                                            avg_Cy, dcx, dcy))

    def replay_chunk(self, lattice, it0, it1, at):
        """The hooks of iterations it0 .. it1-1 in the reference's order
        (printings, outputs, observables, check_stop; run.py:24-54) with the
        observables vectorised: drag and lift of every iteration from the
        device sums, the moving averages by buff.add_chunk (the values mv_avg
        gives), the drag_lift lines in one write.  Returns the iteration
        check_stop ends the run at, or None."""
        from lbm_amd.app.base_app import base_app
        cls = type(self)
        if (cls.observables is not turek.observables or cls.add_buff is not turek.add_buff or
                cls.check_stop is not base_app.check_stop):
            return NotImplemented          # hooks of a subclass: replay them per iteration
        cx, cy = lattice.drag_lift_chunk(self.obstacles[0], self.rho_lbm, self.u_avg, self.D_lbm)
        db, lb = self.drag_buff, self.lift_buff
        prev = (db.obs, lb.obs)
        od, gd = db.add_chunk(cx)
        ol, gl = lb.add_chunk(cy)
        n = it1 - it0
        stop = None
        if self.stop == 'it' and it1 - 1 >= self.it_max:
            stop = max(self.it_max, it0)
        elif self.stop == 'obs' and db.obs_cv and lb.obs_cv:
            # the buffs count the observables calls, 1 at iteration 0
            stop = max(max(db.obs_cv_it, lb.obs_cv_it) - 1, it0)
        m = n if stop is None else stop - it0 + 1
        db.truncate_chunk(m)
        lb.truncate_chunk(m)
        for t in range(m):
            at(it0 + t)
            db.obs, lb.obs = prev if t == 0 else (od[t - 1], ol[t - 1])
            self.printings(it0 + t)
            self.outputs(lattice, it0 + t)
        db.obs, lb.obs = od[m - 1], ol[m - 1]
        dt = self.dt
        self._dl_file.write(''.join(
            '{} {} {} {} {} {} {}\n'.format((it0 + t) * dt, float(cx[t]), float(cy[t]),
                                            float(od[t]), float(ol[t]), float(gd[t]),
                                            float(gl[t])) for t in range(m)))
        if stop is not None:
            self.check_stop(stop)
        return stop

    def finalize(self, lattice):
        f = getattr(self, "_dl_file", None)
        if f is not None:
            f.close()
