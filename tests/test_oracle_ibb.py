"""Pins of the oracle's interpolated bounce-back (nb.py:77-117).

The reference's IBB is the quadratic interpolation of Bouzidi, Firdaouss and
Lallemand (Phys. Fluids 13, 3452, 2001).  Along the link direction s (node x
at s = 0, x - c_q at s = -1, x - 2 c_q at s = -2, wall at s = p), the
population that leaves the wall into x at the next step is, for p < 1/2, the
post-collision g_q carried from s = 2p - 1 (Lagrange interpolation through
s = 0, -1, -2 of plane q); for p >= 1/2, it is the value at s = 0 of the
parabola through (2p - 1, g_up_q(x)), (-1, g_up_qb(x)), (-2, g_up_qb(x - c_q)).
Both interpolations are exact on quadratics, so with those stencil entries
set from a random quadratic phi the result must be phi(2p - 1) (p < 1/2) or
phi(0) (p >= 1/2), every other entry of g must be left alone, and a plane,
offset or direction mistake reads one of the unrelated random values around
the stencil and fails.

test_mutants_fail plants such mistakes in copies of the oracle source, built
into a temporary folder, and checks that this test catches each of them.
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle

C = np.array([[0, 0], [1, 0], [-1, 0], [0, 1], [0, -1], [1, 1], [-1, -1], [-1, 1], [1, -1]])
NS = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7])
P_VALUES = (0.1, 0.3, 0.5, 0.7, 0.95)
NX, NY = 14, 12
X = (6, 5)   # the link's fluid node; x - 2 c_q stays inside for every q


def _ibb_call(lib, bnd, p, gup, g):
    dp = ctypes.c_void_p
    st = lib.oracle_bounce_back_obstacle(
        NX, NY, 1, bnd.shape[0], bnd.ctypes.data_as(dp), p.ctypes.data_as(dp),
        gup.ctypes.data_as(dp), g.ctypes.data_as(dp))
    assert st == 0


def ibb_exactness_errors(lib, seed=0):
    """Largest deviation from the exact quadratic value over every direction
    and distance, and whether any other entry of g changed."""
    rng = np.random.default_rng(seed)
    worst = 0.0
    untouched = True
    for p in P_VALUES:
        for q in range(1, 9):
            qb = NS[q]
            a, b, c2 = rng.uniform(-1.0, 1.0, 3)
            phi = lambda s: a + b * s + c2 * s * s   # noqa: E731
            gup = rng.uniform(-5.0, 5.0, (9, NX, NY))   # unrelated values everywhere
            x0, y0 = X
            cx, cy = C[q]
            if p < 0.5:
                for k in range(3):
                    gup[q, x0 - k * cx, y0 - k * cy] = phi(-k)
                want = phi(2 * p - 1)
            else:
                gup[q, x0, y0] = phi(2 * p - 1)
                gup[qb, x0, y0] = phi(-1.0)
                gup[qb, x0 - cx, y0 - cy] = phi(-2.0)
                want = phi(0.0)
            g0 = rng.uniform(-5.0, 5.0, (9, NX, NY))
            g = g0.copy()
            bnd = np.array([[x0, y0, q]], dtype=np.int64)
            _ibb_call(lib, bnd, np.array([p]), np.ascontiguousarray(gup), g)
            worst = max(worst, abs(g[qb, x0, y0] - want))
            mask = np.ones(g.shape, bool)
            mask[qb, x0, y0] = False
            untouched &= bool(np.array_equal(g[mask], g0[mask]))
    return worst, untouched


def test_ibb_exact_on_quadratics():
    worst, untouched = ibb_exactness_errors(oracle.load())
    assert worst < 1e-13, worst
    assert untouched


def test_ibb_several_links_one_call():
    """Links on separate nodes in one call: each gets its own exact value."""
    rng = np.random.default_rng(7)
    nodes = [(3, 3, 1), (10, 3, 2), (3, 8, 3), (10, 8, 4), (6, 3, 5), (6, 8, 6), (8, 5, 7),
             (4, 6, 8)]
    gup = rng.uniform(-5.0, 5.0, (9, 16, 12))
    ps = np.array([0.1, 0.95, 0.3, 0.7, 0.5, 0.2, 0.8, 0.45])
    want = []
    for (x0, y0, q), p in zip(nodes, ps):
        a, b, c2 = rng.uniform(-1.0, 1.0, 3)
        phi = lambda s: a + b * s + c2 * s * s   # noqa: E731
        cx, cy = C[q]
        if p < 0.5:
            for k in range(3):
                gup[q, x0 - k * cx, y0 - k * cy] = phi(-k)
            want.append(phi(2 * p - 1))
        else:
            gup[q, x0, y0] = phi(2 * p - 1)
            gup[NS[q], x0, y0] = phi(-1.0)
            gup[NS[q], x0 - cx, y0 - cy] = phi(-2.0)
            want.append(phi(0.0))
    g = np.zeros((9, 16, 12))
    oracle.bounce_back_obstacle(True, np.array(nodes, dtype=np.int64), ps, gup, g)
    got = [g[NS[q], x0, y0] for (x0, y0, q) in nodes]
    assert np.allclose(got, want, rtol=0, atol=1e-13)


# Mistakes a restatement of nb.py:85-104 could plausibly make.  Each entry:
# (name, text of the oracle source, replacement).
MUTANTS = [
    ("p<1/2: plane qb instead of q at x - c_q",
     "(1.0 + pp) * (1.0 - pp) * G(g_up, q, im, jm)",
     "(1.0 + pp) * (1.0 - pp) * G(g_up, qb, im, jm)"),
    ("p<1/2: x - c_q and x - 2c_q swapped",
     "(1.0 + pp) * (1.0 - pp) * G(g_up, q, im, jm)\n"
     "                               - p * (1.0 - pp) * G(g_up, q, imm, jmm);",
     "(1.0 + pp) * (1.0 - pp) * G(g_up, q, imm, jmm)\n"
     "                               - p * (1.0 - pp) * G(g_up, q, im, jm);"),
    ("stencil along +c_q instead of -c_q",
     "wrap(boundary[3 * k + 0] + CX[qb], nx, &im) ||\n"
     "                wrap(boundary[3 * k + 1] + CY[qb], ny, &jm) ||\n"
     "                wrap(boundary[3 * k + 0] + 2 * CX[qb], nx, &imm) ||\n"
     "                wrap(boundary[3 * k + 1] + 2 * CY[qb], ny, &jmm)",
     "wrap(boundary[3 * k + 0] - CX[qb], nx, &im) ||\n"
     "                wrap(boundary[3 * k + 1] - CY[qb], ny, &jm) ||\n"
     "                wrap(boundary[3 * k + 0] - 2 * CX[qb], nx, &imm) ||\n"
     "                wrap(boundary[3 * k + 1] - 2 * CY[qb], ny, &jmm)"),
    ("p>=1/2: plane q instead of qb at x - c_q",
     "((1.0 - pp) / (1.0 + pp)) * G(g_up, qb, im, jm)",
     "((1.0 - pp) / (1.0 + pp)) * G(g_up, q, im, jm)"),
    ("p>=1/2: plane q instead of qb at x",
     "((pp - 1.0) / p) * G(g_up, qb, i, j)",
     "((pp - 1.0) / p) * G(g_up, q, i, j)"),
    ("p>=1/2: x - 2c_q instead of x - c_q",
     "((1.0 - pp) / (1.0 + pp)) * G(g_up, qb, im, jm)",
     "((1.0 - pp) / (1.0 + pp)) * G(g_up, qb, imm, jmm)"),
    ("p<1/2: sign of the third term",
     "- p * (1.0 - pp) * G(g_up, q, imm, jmm);",
     "+ p * (1.0 - pp) * G(g_up, q, imm, jmm);"),
    ("result written to plane q instead of qb",
     "G(g, qb, i, j) = p * (pp + 1.0) * G(g_up, q, i, j)",
     "G(g, q, i, j) = p * (pp + 1.0) * G(g_up, q, i, j)"),
]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("name,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_mutants_fail(tmp_path, name, old, new):
    src = open(os.path.join(os.path.dirname(oracle.__file__), "d2q9_oracle.c")).read()
    assert src.count(old) == 1, "mutation site not found: " + name
    path = tmp_path / "mutant.c"
    path.write_text(src.replace(old, new))
    so = tmp_path / "libmutant.so"
    subprocess.check_call(["gcc", "-O1", "-fPIC", "-shared", "-ffp-contract=off",
                           "-std=c99", str(path), "-o", str(so)])
    lib = ctypes.CDLL(str(so))
    lib.oracle_bounce_back_obstacle.restype = ctypes.c_int
    lib.oracle_bounce_back_obstacle.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p] * 4
    worst, untouched = ibb_exactness_errors(lib)
    assert worst > 1e-6 or not untouched, name + " was not caught"
