"""Velocity-magnitude frames (lbm/src/plot/plot.py:9-32)."""
import numpy as np


def plot_norm(lattice, val_min, val_max, output_it, dpi):
    """Write png_dir/u_norm_<output_it>.png of |u| with obstacles masked."""
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    u = lattice.u
    v = np.sqrt(u[0, :, :] ** 2 + u[1, :, :] ** 2)
    v[np.where(lattice.lattice > 0.0)] = -1.0
    vm = np.rot90(np.ma.masked_where((v < 0.0), v))
    plt.clf()
    fig, ax = plt.subplots(figsize=plt.figaspect(vm))
    fig.subplots_adjust(0, 0, 1, 1)
    plt.imshow(vm, cmap='RdBu_r', vmin=val_min * lattice.u_lbm, vmax=val_max * lattice.u_lbm,
               interpolation='spline16')
    plt.axis('off')
    plt.savefig(lattice.png_dir + 'u_norm_' + str(output_it) + '.png', dpi=dpi)
    plt.close('all')
