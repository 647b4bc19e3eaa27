"""Literal restatement of the reference's obstacle set-up -- TEST INFRASTRUCTURE ONLY.

Per-node Python loops exactly as lbm/src/core/lattice.py:290-414 runs them
(get_coords, is_inside, add_obstacle); the checker for the vectorised
lbm_amd.geometry on the reference's own geometries (Turek cylinder, array
and step squares).  Pinned by the reference's test_lattice counts (234 links
at L_lbm=100, 468 at L_lbm=200).
"""
import numpy as np

C = np.array([[0, 0], [1, 0], [-1, 0], [0, 1], [0, -1], [1, 1], [-1, -1], [-1, 1], [1, -1]])
NS = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7])


def get_coords(lat, i, j):
    """lattice.py:379-387"""
    dx = (lat.x_max - lat.x_min) / (lat.nx - 1)
    dy = (lat.y_max - lat.y_min) / (lat.ny - 1)
    x = lat.x_min + i * dx
    y = lat.y_min + j * dy
    return [x, y]


def is_inside(poly, pt):
    """lattice.py:391-414"""
    j = len(poly) - 1
    odd_nodes = False
    for i in range(len(poly)):
        if (((poly[i, 1] < pt[1] and poly[j, 1] >= pt[1]) or
             (poly[j, 1] < pt[1] and poly[i, 1] >= pt[1])) and
                (poly[i, 0] < pt[0] or poly[j, 0] < pt[0])):
            slope = (poly[j, 0] - poly[i, 0]) / (poly[j, 1] - poly[i, 1])
            if (poly[i, 0] + (pt[1] - poly[i, 1]) * slope) < pt[0]:
                odd_nodes = not odd_nodes
        j = i
    return odd_nodes


def add_obstacle(lat, obstacle):
    """lattice.py:290-375; returns (area, boundary, ibb)."""
    tag = obstacle.tag
    polygon = obstacle.polygon
    bnds = [np.amin(polygon[:, 0]), np.amax(polygon[:, 0]),
            np.amin(polygon[:, 1]), np.amax(polygon[:, 1])]
    obs = []
    for i in range(lat.nx):
        for j in range(lat.ny):
            pt = get_coords(lat, i, j)
            if pt[0] > bnds[0] and pt[0] < bnds[1] and pt[1] > bnds[2] and pt[1] < bnds[3]:
                if is_inside(polygon, pt):
                    lat.lattice[i, j] = tag
                    obs.append([i, j])
    bnd = []
    for (i, j) in obs:
        for q in range(1, 9):
            qb = NS[q]
            ii = i + C[q, 0]
            jj = j + C[q, 1]
            if ii > lat.nx - 1 or jj > lat.ny - 1:
                continue
            if not lat.lattice[ii, jj]:
                bnd.append([ii, jj, qb])
    bnd = np.unique(np.array(bnd, dtype=int).reshape(-1, 3), axis=0)
    ibb = []
    if lat.IBB:
        for k in range(len(bnd)):
            i, j, q = bnd[k]
            pt = get_coords(lat, i, j)
            x = polygon[:, 0] - pt[0]
            y = polygon[:, 1] - pt[1]
            dist = np.sqrt(np.square(x) + np.square(y))
            mpt = np.argmin(dist)
            ibb.append(dist[mpt] / (lat.dx * np.linalg.norm(C[q])))
    area = 0.0
    for i in range(lat.nx):
        for j in range(lat.ny):
            if lat.lattice[i, j] == tag:
                area += lat.dx ** 2
    return area, bnd, np.array(ibb, dtype=float)
