"""Obstacle polygons: cylinders, squares and random Bezier shapes.

Host-side setup code (runs once per obstacle), restating
lbm/src/utils/shapes.py:253-466.  The polygon feeds the link extraction of
``lbm_amd.geometry``, whose output (links and IBB distances) must match the
reference's, so the numerical steps are the reference's: the same NumPy
primitives on the same operands in the same order.  Images are optional
(``write_png``), the CSV is written as the reference writes it (shapes.py:178-190).
"""
import math
import os

import numpy as np
import scipy.special


def _distance(p1, p2):
    return np.sqrt((p1[0] - p2[0]) ** 2 + (p1[1] - p2[1]) ** 2)


def cylinder_points(n_pts):
    """Control points on the circle of diameter 1 (shapes.py:259-270)."""
    if n_pts < 4:
        raise ValueError("a cylinder needs at least 4 points")
    ang = 2.0 * math.pi / n_pts
    return np.array([[0.5 * math.cos(float(i) * ang), 0.5 * math.sin(float(i) * ang)]
                     for i in range(n_pts)])


def square_points(n_pts):
    """Corners of the unit square centred on 0 (shapes.py:274-287)."""
    if n_pts != 4:
        raise ValueError("a square needs n_pts = 4")
    return 0.5 * np.array([[1.0, 1.0], [-1.0, 1.0], [-1.0, -1.0], [1.0, -1.0]])


def _dedupe(pts, tol=1.0e-8):
    """Drop points closer than tol to an earlier kept point (shapes.py:292-321)."""
    n = len(pts)
    removed = np.zeros(n, dtype=bool)
    idx = np.arange(n)
    for i in range(n):
        if removed[i]:
            continue
        d = np.sqrt((pts[i, 0] - pts[:, 0]) ** 2 + (pts[i, 1] - pts[:, 1]) ** 2)
        removed |= (d < tol) & ~removed & (idx != i)
    return pts[~removed]


def _ccw_order(pts):
    """Counter-clockwise order around the centroid (shapes.py:329-338)."""
    t = pts - np.mean(pts, axis=0)
    return np.arctan2(t[:, 1], t[:, 0]).argsort()


def _bezier(ctrl, n):
    """Cubic Bezier sampled at n points (shapes.py:342-362)."""
    t = np.linspace(0, 1, n)
    out = np.zeros((n, 2))
    m = len(ctrl) - 1
    for k in range(len(ctrl)):
        b = scipy.special.binom(m, k) * (t ** k) * ((1.0 - t) ** (m - k))
        out += np.outer(b, ctrl[k])
    return out


def _segment(p1, p2, d1, d2, db1, db2, r1, r2, e1, e2, n):
    """Bezier segment between two control points (shapes.py:366-403)."""
    if n == 0:
        return np.vstack([p1, p2])
    ctrl = np.zeros((4, 2))
    ctrl[0, :] = p1[:]
    ctrl[3, :] = p2[:]
    ctrl[1, :] = p1 + e1 * (r1 * d1) + (1.0 - e1) * (r1 * db1)
    ctrl[2, :] = p2 + e2 * (-r2 * d2) + (1.0 - e2) * (r2 * db2)
    return _bezier(ctrl, n)


class Shape:
    """A closed shape: control points, radii, edginess and the sampled curve."""

    def __init__(self, name, position, control_pts, n_sampling_pts, radius, edgy, output_dir):
        self.name = name
        self.position = position
        self.control_pts = control_pts
        self.n_control_pts = len(control_pts)
        self.n_sampling_pts = n_sampling_pts
        self.radius = radius
        self.edgy = edgy
        self.output_dir = output_dir
        self.curve_pts = np.array([])

    def build(self):
        """Sample the closed curve (shapes.py:57-127)."""
        n = self.n_control_pts
        self.control_pts -= np.mean(self.control_pts, axis=0)
        order = _ccw_order(self.control_pts)
        pts = np.array(self.control_pts)[order, :]
        rad = np.array(self.radius)[order]
        edg = np.array(self.edgy)[order]
        delta = np.zeros([n, 2])
        delta_b = np.zeros([n, 2])
        radii = np.zeros([n, 2])
        for i in range(n):
            pm, pc, pn = pts[i - 1, :], pts[i, :], pts[(i + 1) % n, :]
            d = pn - pm
            delta[i, :] = d / np.linalg.norm(d)
            delta_b[i, :] = 0.5 * (pm + pn) - pc
            radii[i, 0] = 0.5 * _distance(pm, pc) * rad[i]
            radii[i, 1] = 0.5 * _distance(pc, pn) * rad[i]
        pieces = []
        for i in range(n):
            j = (i + 1) % n
            smpl = math.ceil(self.n_sampling_pts * math.sqrt(_distance(pts[i, :], pts[j, :])))
            pieces.append(_segment(pts[i, :], pts[j, :], delta[i, :], delta[j, :],
                                   delta_b[i, :], delta_b[j, :], radii[i, 1], radii[j, 0],
                                   edg[i], edg[j], smpl))
        curve = np.concatenate(pieces)
        x, y = curve.T
        curve_pts = _dedupe(np.column_stack((x, y, np.zeros(x.size))))
        center = np.mean(curve_pts, axis=0)
        curve_pts -= center
        self.control_pts[:, 0:2] -= center[0:2]
        self.control_pts[:, 0:2] += self.position[0:2]
        curve_pts[:, 0:2] += self.position[0:2]
        self.curve_pts = curve_pts

    def write_csv(self):
        """Control points, radii and edginess (shapes.py:178-190 format)."""
        os.makedirs(self.output_dir, exist_ok=True)
        with open(os.path.join(self.output_dir, self.name + ".csv"), "w") as f:
            f.write("{} {}\n".format(self.n_control_pts, self.n_sampling_pts))
            for i in range(self.n_control_pts):
                f.write("{} {} {} {}\n".format(self.control_pts[i, 0], self.control_pts[i, 1],
                                               self.radius[i], self.edgy[i]))

    def write_png(self, size):
        """Picture of the shape (shapes.py:131-174 look, without trimming)."""
        import matplotlib
        matplotlib.use("Agg")
        import matplotlib.pyplot as plt
        fig = plt.figure()
        plt.xlim([-size, size])
        plt.ylim([-size, size])
        plt.axis("off")
        plt.gca().set_aspect("equal", adjustable="box")
        plt.fill([-size, size, size, -size], [-size, -size, size, size],
                 color=(0.784, 0.773, 0.741), linewidth=2.5, zorder=0)
        plt.fill(self.curve_pts[:, 0], self.curve_pts[:, 1], "black", linewidth=0, zorder=1)
        plt.savefig(os.path.join(self.output_dir, self.name + ".png"), dpi=200)
        plt.close(fig)


def generate_shape(n_pts, position, shape_type, shape_size, shape_name, n_sampling_pts,
                   output_dir, write_png=False):
    """Build an obstacle shape (shapes.py:417-466)."""
    if shape_type == "cylinder":
        radius = 0.5 * np.ones((n_pts))
        edgy = 1.0 * np.ones((n_pts))
        ctrl = cylinder_points(n_pts)
    elif shape_type == "square":
        radius = np.zeros((n_pts))
        edgy = np.ones((n_pts))
        ctrl = square_points(n_pts)
    elif shape_type == "random":
        radius = np.random.uniform(low=0.8, high=1.0, size=n_pts)
        edgy = np.random.uniform(low=0.45, high=0.55, size=n_pts)
        ctrl = np.random.rand(n_pts, 2)
    else:
        raise ValueError('shape_type must be "cylinder", "square" or "random"')
    ctrl[:, :] *= shape_size
    s = Shape(shape_name, position, ctrl, n_sampling_pts, radius, edgy, output_dir)
    s.build()
    if write_png:
        s.write_png(shape_size)
    s.write_csv()
    return s
