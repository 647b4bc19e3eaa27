"""Pins of the CPU oracle (oracle/d2q9_oracle.c via oracle/oracle.py).

Reference golden data: the reference's own cavity and Poiseuille tests.
Closed forms and invariants, each chosen so that a plausible slip (a wrong
weight, sign, index, coefficient pairing or streaming direction) fails:
equilibrium moments, collision invariants, streaming along c_q, the rest
state as a fixed point, Zou-He walls imposing their velocity/density, corners
at rest, IBB preserving a uniform field for every distance p (and reducing to
plain bounce-back at p = 1/2), drag of a single link and of a symmetric body.
The IBB stencil itself (planes, offsets, direction) is pinned by the
quadratic-exactness checks of test_oracle_ibb.py.
"""
import os

import numpy as np
import pytest

from oracle import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")
C = np.array([[0, 0], [1, 0], [-1, 0], [0, 1], [0, -1], [1, 1], [-1, -1], [-1, 1], [1, -1]])
W = np.array([4. / 9.] + [1. / 9.] * 4 + [1. / 36.] * 4)
NS = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7])


def _rand_g(rng, nx, ny, amp=0.2):
    return np.ascontiguousarray(W[:, None, None] * (1.0 + amp * rng.standard_normal((9, nx, ny))))


# ----------------------------------------------------- reference goldens --
def test_reference_cavity_profiles():
    """lbm/tst/cavity/test_cavity.py: default cavity, 10 001 iterations.

    The reference asserts rows 10 and 50 of both profiles to 1e-6.  The whole
    200-value profiles of its golden files are reproduced to 3.2e-6 (they were
    written by the reference with fastmath Numba and BLAS sums); 5e-6 here."""
    from tests.cases import _Quiet
    from lbm_amd.app.cavity import cavity
    app = cavity()
    app.__class__ = type("q_cavity", (_Quiet, cavity), {})
    app.output_freq = 10000
    lat = oracle.OracleLattice(app, results_dir=_Quiet.results_dir)
    oracle.run(lat, app)
    vx, uy = app.line_fields(lat)
    assert abs(vx[10] - 0.12493089684236539) < 1.0e-6
    assert abs(vx[50] - 0.05295104908939561) < 1.0e-6
    assert abs(uy[10] + 0.05968571489630510) < 1.0e-6
    assert abs(uy[50] + 0.19792323493599165) < 1.0e-6
    ref_vx = np.loadtxt(os.path.join(GOLD, "cavity_vx_ref"))
    ref_uy = np.loadtxt(os.path.join(GOLD, "cavity_uy_ref"))
    assert np.abs(vx - ref_vx[:, 1]).max() < 5e-6
    assert np.abs(uy - ref_uy[:, 1]).max() < 5e-6


def test_reference_poiseuille():
    """lbm/tst/poiseuille/test_poiseuille.py: L1 error < 1e-3 against the parabola."""
    from tests.cases import _Quiet
    from lbm_amd.app.poiseuille import poiseuille
    app = poiseuille()
    app.__class__ = type("q_pois", (_Quiet, poiseuille), {})
    lat = oracle.OracleLattice(app, results_dir=_Quiet.results_dir)
    oracle.run(lat, app)
    assert app.compute_error(lat) < 1.0e-3


# ------------------------------------------------ closed forms/invariants --
def test_macro_of_weights_is_rest():
    g = np.ascontiguousarray(np.broadcast_to(W[:, None, None], (9, 5, 4)))
    rho, u = np.zeros((5, 4)), np.zeros((2, 5, 4))
    oracle.macro(g, rho, u)
    assert np.allclose(rho, 1.0, rtol=0, atol=1e-15)
    assert np.abs(u).max() < 1e-16


def test_equilibrium_moments():
    rng = np.random.default_rng(0)
    rho = 1.0 + 0.05 * rng.standard_normal((7, 6))
    u = 0.08 * rng.standard_normal((2, 7, 6))
    geq = np.zeros((9, 7, 6))
    oracle.equilibrium(u, rho, geq)
    assert np.allclose(geq.sum(axis=0), rho, rtol=1e-14, atol=0)
    mom = np.tensordot(C.T.astype(float), geq, axes=(1, 0))
    assert np.allclose(mom, rho * u, rtol=0, atol=1e-15)
    # second moment: sum c_a c_b g_eq = rho (u_a u_b + delta_ab / 3)
    pxx = np.tensordot((C[:, 0] * C[:, 0]).astype(float), geq, axes=(0, 0))
    assert np.allclose(pxx, rho * (u[0] ** 2 + 1.0 / 3.0), rtol=1e-13, atol=1e-15)


def test_collision_conserves_mass_and_momentum():
    rng = np.random.default_rng(1)
    nx, ny = 9, 7
    g = _rand_g(rng, nx, ny)
    rho, u = np.zeros((nx, ny)), np.zeros((2, nx, ny))
    oracle.macro(g, rho, u)
    geq = np.zeros_like(g)
    oracle.equilibrium(u, rho, geq)
    gup = np.zeros_like(g)
    g0 = g.copy()
    oracle.col_str(g, geq, gup, 1.0 / 0.8, 1.0 / 1.7)
    assert np.allclose(gup.sum(axis=0), g0.sum(axis=0), rtol=1e-14, atol=0)
    for d in range(2):
        a = np.tensordot(C[:, d].astype(float), gup, axes=(0, 0))
        b = np.tensordot(C[:, d].astype(float), g0, axes=(0, 0))
        assert np.allclose(a, b, rtol=0, atol=1e-15)
    # non-conserved modes do relax: the result differs from a plain copy
    assert not np.allclose(gup, g0)


def test_streaming_moves_along_c():
    nx, ny = 8, 7
    for q in range(1, 9):
        g = np.zeros((9, nx, ny))
        geq = np.zeros_like(g)
        geq[q, 4, 3] = 1.0
        gup = np.zeros_like(g)
        # om_p = om_m = 1: g_up = g_eq exactly
        oracle.col_str(g, geq, gup, 1.0, 1.0)
        assert gup[q, 4, 3] == 1.0
        assert g[q, 4 + C[q, 0], 3 + C[q, 1]] == 1.0
        assert g[q].sum() == 1.0


def test_rest_state_is_a_fixed_point():
    nx, ny = 10, 8
    g = np.ascontiguousarray(np.broadcast_to(W[:, None, None], (9, nx, ny)))
    rho, u = np.ones((nx, ny)), np.zeros((2, nx, ny))
    geq, gup = np.zeros_like(g), np.zeros_like(g)
    oracle.macro(g, rho, u)
    oracle.equilibrium(u, rho, geq)
    oracle.col_str(g, geq, gup, 1.0 / 0.7, 1.0 / 1.3)
    zx, zy = np.zeros((2, nx)), np.zeros((2, ny))
    oracle.zou_he("bottom", u, rho, g, side=zx)
    oracle.zou_he("left", u, rho, g, side=zy)
    oracle.zou_he("top", u, rho, g, side=zx)
    oracle.zou_he("right_pressure", u, rho, g, side=zy, rho_right=np.ones(ny))
    for c in ("bottom_left", "top_left", "top_right", "bottom_right"):
        oracle.zou_he(c, u, rho, g)
    corner = np.zeros((nx, ny), bool)
    corner[[0, 0, -1, -1], [0, -1, -1, 0]] = True
    assert np.allclose(g[:, ~corner], W[:, None], rtol=0, atol=1e-15)
    # corners (nb.py:251-344): buried diagonals zeroed, their weight moved to g0
    for (i, j, buried) in ((0, 0, (7, 8)), (0, -1, (5, 6)), (-1, -1, (7, 8)), (-1, 0, (5, 6))):
        assert g[buried[0], i, j] == 0.0 and g[buried[1], i, j] == 0.0
        assert np.isclose(g[0, i, j], 4. / 9. + 2. / 36., rtol=1e-14)


@pytest.mark.parametrize("side", ["left", "right", "top", "bottom"])
def test_zou_he_wall_imposes_velocity(side):
    rng = np.random.default_rng(2)
    nx, ny = 9, 8
    g = _rand_g(rng, nx, ny, 0.1)
    rho, u = np.ones((nx, ny)), np.zeros((2, nx, ny))
    n = ny if side in ("left", "right") else nx
    vel = 0.05 * rng.standard_normal((2, n))
    oracle.zou_he(side, u, rho, g, side=np.ascontiguousarray(vel))
    if side == "left":
        sl = (slice(None), 0, slice(1, ny - 1))
    elif side == "right":
        sl = (slice(None), nx - 1, slice(1, ny - 1))
    elif side == "top":
        sl = (slice(None), slice(1, nx - 1), ny - 1)
    else:
        sl = (slice(None), slice(1, nx - 1), 0)
    gw = g[sl]
    r = gw.sum(axis=0)
    v = vel[:, 1:-1]
    mx = np.tensordot(C[:, 0].astype(float), gw, axes=(0, 0))
    my = np.tensordot(C[:, 1].astype(float), gw, axes=(0, 0))
    assert np.allclose(r, rho[sl[1:]], rtol=1e-14)
    assert np.allclose(mx, r * v[0], rtol=0, atol=1e-15)
    assert np.allclose(my, r * v[1], rtol=0, atol=1e-15)


def test_zou_he_pressure_outlet_imposes_density():
    rng = np.random.default_rng(3)
    nx, ny = 9, 8
    g = _rand_g(rng, nx, ny, 0.1)
    rho, u = np.ones((nx, ny)), np.zeros((2, nx, ny))
    uy = np.zeros((2, ny))
    uy[1] = 0.02 * rng.standard_normal(ny)
    rr = 1.0 + 0.01 * rng.standard_normal(ny)
    oracle.zou_he("right_pressure", u, rho, g, side=uy, rho_right=rr)
    gw = g[:, nx - 1, 1:ny - 1]
    r = gw.sum(axis=0)
    assert np.allclose(r, rr[1:-1], rtol=1e-14)
    my = np.tensordot(C[:, 1].astype(float), gw, axes=(0, 0))
    mx = np.tensordot(C[:, 0].astype(float), gw, axes=(0, 0))
    assert np.allclose(my, r * uy[1, 1:-1], atol=1e-15)
    assert np.allclose(mx, r * u[0, nx - 1, 1:-1], atol=1e-15)


@pytest.mark.parametrize("corner,cell", [("bottom_left", (0, 0)), ("top_left", (0, -1)),
                                         ("top_right", (-1, -1)), ("bottom_right", (-1, 0))])
def test_corner_at_rest(corner, cell):
    """Zero wall velocities: the corner node carries the neighbour's density
    and no momentum; the two buried diagonals are zero."""
    rng = np.random.default_rng(4)
    nx, ny = 8, 7
    g = _rand_g(rng, nx, ny, 0.1)
    rho, u = np.ones((nx, ny)), np.zeros((2, nx, ny))
    zx, zy = np.zeros((2, nx)), np.zeros((2, ny))
    for s, z in (("bottom", zx), ("left", zy), ("right", zy), ("top", zx)):
        oracle.zou_he(s, u, rho, g, side=z)
    oracle.zou_he(corner, u, rho, g)
    i, j = cell
    gc = g[:, i, j]
    ni = 1 if i == 0 else nx - 2
    assert np.isclose(gc.sum(), rho[ni, j], rtol=1e-14)
    assert abs(np.dot(C[:, 0], gc)) < 1e-16 and abs(np.dot(C[:, 1], gc)) < 1e-16
    assert rho[i, j] == rho[ni, j]


@pytest.mark.parametrize("p", [0.01, 0.2, 0.4999, 0.5, 0.73, 1.0, 1.12])
def test_ibb_preserves_uniform_field(p):
    nx, ny = 12, 10
    gup = np.full((9, nx, ny), 0.037)
    g = np.zeros((9, nx, ny))
    bnd = np.array([[5, 4, q] for q in range(1, 9)], dtype=np.int64)
    oracle.bounce_back_obstacle(True, bnd, np.full(8, p), gup, g)
    for q in range(1, 9):
        assert np.isclose(g[NS[q], 5, 4], 0.037, rtol=1e-14)


def test_ibb_half_distance_is_bounce_back():
    rng = np.random.default_rng(5)
    nx, ny = 12, 10
    gup = _rand_g(rng, nx, ny)
    bnd = np.array([[5, 4, q] for q in range(1, 9)], dtype=np.int64)
    g1, g2 = np.zeros_like(gup), np.zeros_like(gup)
    oracle.bounce_back_obstacle(True, bnd, np.full(8, 0.5), gup, g1)
    oracle.bounce_back_obstacle(False, bnd, None, gup, g2)
    assert np.array_equal(g1, g2)
    for q in range(1, 9):
        assert g2[NS[q], 5, 4] == gup[q, 5, 4]


def test_drag_single_link_and_symmetric_body():
    nx, ny = 12, 10
    gup = np.zeros((9, nx, ny))
    g = np.zeros((9, nx, ny))
    gup[5, 3, 3] = 0.25
    g[6, 3, 3] = 0.5
    assert oracle.drag_sums(np.array([[3, 3, 5]]), gup, g) == (0.75, 0.75)
    # all links around a 2x2 block in a uniform field: the forces cancel
    solid = np.zeros((nx, ny), bool)
    solid[5:7, 4:6] = True
    bnd = [[i, j, q] for i in range(1, nx - 1) for j in range(1, ny - 1) if not solid[i, j]
           for q in range(1, 9) if solid[i + C[q, 0], j + C[q, 1]]]
    bnd = np.array(bnd, dtype=np.int64)
    gu = np.ascontiguousarray(np.broadcast_to(W[:, None, None], (9, nx, ny)))
    fx, fy = oracle.drag_sums(bnd, gu, gu.copy())
    assert abs(fx) < 1e-15 and abs(fy) < 1e-15
    cx, cy = oracle.drag_lift(bnd, gu, gu.copy(), 1.0, 0.1, 2.0)
    assert abs(cx) < 1e-12 and abs(cy) < 1e-12
