"""Base app hooks (lbm/src/app/base_app.py:10-69)."""
from lbm_amd.shapes import generate_shape


class base_app():
    output_freq = 500
    output_it = 0

    def set_inlets(self, lattice, it):
        pass

    def observables(self, lattice, it):
        pass

    def finalize(self, lattice):
        pass

    def outputs(self, lattice, it):
        pass

    def add_obstacles(self, lattice, obstacles):
        """Shape, tag and links of each obstacle (base_app.py:27-38)."""
        for i in range(len(obstacles)):
            obs = obstacles[i]
            shape = generate_shape(obs.n_pts, obs.pos, obs.type, obs.size, obs.name,
                                   obs.n_spts, lattice.output_dir)
            obstacles[i].set_polygon(shape.curve_pts)
            obstacles[i].set_tag(i + 1)
            area, bnd, ibb = lattice.add_obstacle(obstacles[i])
            obstacles[i].fill(area, bnd, ibb)

    def printings(self, it):
        if not getattr(self, "verbose", True):
            return
        if self.stop == 'it':
            print('# it = ' + str(it) + ' / ' + str(self.it_max), end='\r')
        if self.stop == 'obs':
            str_d = "{:10.6f}".format(self.drag_buff.obs)
            str_l = "{:10.6f}".format(self.lift_buff.obs)
            print('# it = ' + str(it) + ', avg drag =' + str_d + ', avg lift =' + str_l, end='\r')

    def check_stop(self, it):
        compute = True
        if self.stop == 'it':
            if it >= self.it_max:
                compute = False
                if getattr(self, "verbose", True):
                    print('\n')
                    print('# Computation ended: it>it_max')
        if self.stop == 'obs':
            if self.drag_buff.obs_cv and self.lift_buff.obs_cv:
                compute = False
                if getattr(self, "verbose", True):
                    print('\n')
                    print('# Computation ended: converged')
        return compute
