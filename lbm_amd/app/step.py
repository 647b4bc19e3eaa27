"""Channel with two steps (lbm/src/app/step.py:13-139)."""
import math

from lbm_amd.app._channel import channel_app
from lbm_amd.obstacle import obstacle


class step(channel_app):
    def __init__(self):
        self.name = 'step'
        self.Re_lbm = 500.0
        self.L_lbm = 150
        self.u_lbm = 0.05
        self.rho_lbm = 1.0
        self.t_max = 15.0
        self.x_min = -1.0
        self.x_max = 15.0
        self.y_min = -1.0
        self.y_max = 1.0
        self.IBB = False
        self.stop = 'it'
        self.obs_cv_ct = 1.0e-3
        self.obs_cv_nb = 1000
        self.output_freq = 500
        self.output_it = 0
        self.dpi = 200
        self.compute_lbm_parameters()
        self.obstacles = [obstacle('square1', 4, 100, 'square', 1.5, [1.5, 1.0]),
                          obstacle('square2', 4, 100, 'square', 1.5, [4.0, -1.0])]

    def compute_lbm_parameters(self):
        self.Cs = 1.0 / math.sqrt(3.0)
        self.ny = self.L_lbm
        self.u_avg = 2.0 * self.u_lbm / 3.0
        self.r_cyl = 0.5
        self.D_lbm = math.floor(self.ny * self.r_cyl / (self.y_max - self.y_min))
        self.nu_lbm = self.u_avg * self.L_lbm / self.Re_lbm
        self.tau_lbm = 0.5 + self.nu_lbm / (self.Cs ** 2)
        self.dt = self.Re_lbm * self.nu_lbm / self.L_lbm ** 2
        self.dx = (self.y_max - self.y_min) / self.ny
        self.dy = self.dx
        self.nx = math.floor(self.ny * (self.x_max - self.x_min) / (self.y_max - self.y_min))
        self.it_max = math.floor(self.t_max / self.dt)
        self.sigma = math.floor(10 * self.nx)

    def initialize(self, lattice):
        self.init_obstacle_fields(lattice)
        lattice.equilibrium()
        lattice.g = lattice.g_eq.copy()
