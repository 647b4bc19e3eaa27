"""Shared pieces of the channel apps (turek, poiseuille, array, step)."""
import math

import numpy as np

from lbm_amd.app.base_app import base_app
from lbm_amd.plot import plot_norm


class channel_app(base_app):
    """Parabolic Zou-He inlet ramped in time, Zou-He pressure outlet, no-slip
    top and bottom walls (turek.py:94-125 and its siblings)."""

    def poiseuille(self, pt):
        """Unit-amplitude parabola on [y_min, y_max] (turek.py:165-173)."""
        y = pt[1]
        H = self.y_max - self.y_min
        u = np.zeros(2)
        u[0] = 4.0 * (self.y_max - y) * (y - self.y_min) / H ** 2
        return u

    def _inlet_profile(self, lattice):
        prof = getattr(self, "_prof", None)
        if prof is None or prof.shape[1] != lattice.ny:
            prof = np.zeros((2, lattice.ny))
            for j in range(lattice.ny):
                prof[:, j] = self.poiseuille(lattice.get_coords(0, j))
            self._prof = prof
        return prof

    def ramp(self, it):
        """Inlet ramp 1 - exp(-it^2 / (2 sigma^2)) (turek.py:99-100)."""
        val = it
        return (1.0 - math.exp(-val ** 2 / (2.0 * self.sigma ** 2)))

    def separable_inlets(self, lattice):
        """set_inlets(it) as amplitude(it) x profile (the fused run's device
        tables): u_left = (ramp(it) u_lbm) x the parabola, rho_right =
        1 x rho_lbm, the other sides zero."""
        prof = self._inlet_profile(lattice)
        rho = np.full(lattice.ny, float(self.rho_lbm))

        def amp(it):
            return (0.0, self.ramp(it) * self.u_lbm, 0.0, 0.0, 1.0)
        return [None, prof, None, None, rho], amp

    def set_inlets(self, lattice, it):
        # Same values as the reference's per-node loop: (ret*u_lbm) * profile_j.
        ret = self.ramp(it)
        lattice.u_left[:, :] = ret * self.u_lbm * self._inlet_profile(lattice)
        lattice.u_top[0, :] = 0.0
        lattice.u_bot[0, :] = 0.0
        lattice.u_right[1, :] = 0.0
        lattice.rho_right[:] = self.rho_lbm

    def set_bc(self, lattice):
        for obs in getattr(self, "obstacles", []):
            lattice.bounce_back_obstacle(obs)
        lattice.zou_he_bottom_wall_velocity()
        lattice.zou_he_left_wall_velocity()
        lattice.zou_he_top_wall_velocity()
        lattice.zou_he_right_wall_pressure()
        lattice.zou_he_bottom_left_corner()
        lattice.zou_he_top_left_corner()
        lattice.zou_he_top_right_corner()
        lattice.zou_he_bottom_right_corner()

    def outputs(self, lattice, it):
        if it % self.output_freq != 0:
            return
        if getattr(self, "write_png", True):
            plot_norm(lattice, 0.0, 1.5, self.output_it, self.dpi)
        self.output_it += 1

    def init_obstacle_fields(self, lattice):
        if getattr(self, "obstacles", None):
            self.add_obstacles(lattice, self.obstacles)
        self.set_inlets(lattice, 0)
        if getattr(self, "obstacles", None):
            lattice.u[:, np.where(lattice.lattice > 0.0)] = 0.0
        lattice.rho *= self.rho_lbm
        lattice.generate_image(getattr(self, "obstacles", []))
