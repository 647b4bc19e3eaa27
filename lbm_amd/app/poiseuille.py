"""Plane Poiseuille flow (lbm/src/app/poiseuille.py:13-152)."""
import math

import numpy as np

from lbm_amd.app._channel import channel_app


class poiseuille(channel_app):
    def __init__(self):
        self.name = 'poiseuille'
        self.Re_lbm = 100.0
        self.L_lbm = 50
        self.u_lbm = 0.1
        self.rho_lbm = 1.0
        self.t_max = 15.0
        self.x_min = -0.2
        self.x_max = 1.0
        self.y_min = -0.2
        self.y_max = 0.2
        self.stop = 'it'
        self.output_freq = 50000
        self.output_it = 0
        self.dpi = 200
        self.compute_lbm_parameters()

    def compute_lbm_parameters(self):
        self.Cs = 1.0 / math.sqrt(3.0)
        self.ny = self.L_lbm
        self.u_avg = 2.0 * self.u_lbm / 3.0
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

    def finalize(self, lattice):
        self.compute_error(lattice)

    def compute_error(self, lattice):
        """L1 error of u_x at mid-channel against the parabola (poiseuille.py:128-152)."""
        u_error = np.zeros((3, self.ny))
        nx = math.floor(self.nx / 2)
        l1_error = 0.0
        u_all = lattice.u
        for j in range(self.ny):
            pt = lattice.get_coords(nx, j)
            u_ex = self.poiseuille(pt)
            u = u_all[:, nx, j]
            u_error[0, j] = u[0] / self.u_lbm
            u_error[1, j] = u_ex[0]
            u_error[2, j] = abs(u_error[0, j] - u_error[1, j])
            l1_error += u_error[2, j] * self.dx
        with open(lattice.output_dir + 'poiseuille', 'w') as f:
            for j in range(self.ny):
                f.write('{} {} {}\n'.format(j * self.dx, u_error[0, j], u_error[1, j]))
        return l1_error
