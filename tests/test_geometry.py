"""Obstacle set-up (host): the vectorised lbm_amd.geometry against the
reference's test_lattice counts, against a literal restatement of the
reference's per-node loops (oracle/geometry.py), and the shape builder
against the reference's saved control points."""
import os

import numpy as np
import pytest

from lbm_amd import geometry
from lbm_amd.shapes import generate_shape
from oracle import geometry as literal
from tests.cases import _TMP

GOLD = os.path.join(os.path.dirname(__file__), "golden")


class _Lat:
    def __init__(self, app):
        for k in ("nx", "ny", "x_min", "x_max", "y_min", "y_max", "dx"):
            setattr(self, k, getattr(app, k))
        self.IBB = getattr(app, "IBB", False)
        self.lattice = np.zeros((self.nx, self.ny))


def _shape(o):
    return generate_shape(o.n_pts, o.pos, o.type, o.size, o.name, o.n_spts, _TMP + "/")


@pytest.mark.parametrize("L,count", [(100, 234), (200, 468)])
def test_reference_link_counts(L, count):
    """lbm/tst/lattice/test_lattice.py:27,36."""
    from lbm_amd.app.turek import turek
    app = turek()
    app.L_lbm = L
    app.compute_lbm_parameters()
    lat = _Lat(app)
    o = app.obstacles[0]
    o.set_polygon(_shape(o).curve_pts)
    o.set_tag(1)
    area, bnd, ibb = geometry.add_obstacle(lat, o, verbose=False)
    assert bnd.shape[0] == count
    assert ibb.shape[0] == count


@pytest.mark.parametrize("app_name", ["turek", "array", "step"])
def test_vectorised_matches_literal_loops(app_name):
    from lbm_amd.app.app import app_factory
    app = app_factory.create(app_name)
    if app_name == "turek":
        app.L_lbm = 60
        app.compute_lbm_parameters()
    if app_name == "array":
        app.L_lbm = 80
        app.compute_lbm_parameters()
    if app_name == "step":
        app.L_lbm = 40
        app.compute_lbm_parameters()
    l1, l2 = _Lat(app), _Lat(app)
    for k, o in enumerate(app.obstacles):
        o.set_polygon(_shape(o).curve_pts)
        o.set_tag(k + 1)
        a1, b1, i1 = geometry.add_obstacle(l1, o, verbose=False)
        a2, b2, i2 = literal.add_obstacle(l2, o)
        assert np.array_equal(b1, b2)
        assert np.array_equal(i1, i2)
        assert a1 == a2
    assert np.array_equal(l1.lattice, l2.lattice)


def test_step_links_wrap_like_numpy():
    """The step app's lower square touches the bottom wall: its links towards
    y = -1 are kept with j = -1 (NumPy index wrap), as the reference keeps them."""
    from lbm_amd.app.step import step
    app = step()
    app.L_lbm = 40
    app.compute_lbm_parameters()
    lat = _Lat(app)
    o = app.obstacles[1]
    o.set_polygon(_shape(o).curve_pts)
    o.set_tag(2)
    _, bnd, _ = geometry.add_obstacle(lat, o, verbose=False)
    assert (bnd[:, 1] == -1).any()


@pytest.mark.parametrize("app_name,which,csv,tol", [
    ("turek", 0, "turek_cylinder.csv", 1e-16),
    ("step", 0, "step_square1.csv", 0.0),
    ("step", 1, "step_square2.csv", 0.0),
    ("array", 7, "array_square7.csv", 1e-16),
])
def test_shape_control_points(app_name, which, csv, tol):
    """Control points after centring on the sampled curve (shapes.py:120-127);
    the centring subtracts the mean of the Bezier samples, so this pins the
    sampled polygon's centroid to the last bits."""
    from lbm_amd.app.app import app_factory
    app = app_factory.create(app_name)
    o = app.obstacles[which]
    s = _shape(o)
    ref = np.loadtxt(os.path.join(GOLD, csv), skiprows=1)
    assert np.abs(s.control_pts - ref[:, :2]).max() <= tol
    assert np.array_equal(s.radius, ref[:, 2]) and np.array_equal(s.edgy, ref[:, 3])
