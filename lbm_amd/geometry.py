"""Obstacle rasterisation, boundary links and IBB distances (vectorised).

Host-side setup, run once per obstacle.  Same results as the reference's
``lattice.add_obstacle`` / ``get_coords`` / ``is_inside``
(lbm/src/core/lattice.py:290-414), whose per-node Python loops take hours on
an 8192x2048 lattice; here every step is a NumPy array operation:

* inside test: the even-odd ray-crossing rule of lattice.py:391-414, edge by
  edge over all nodes of the polygon's bounding box at once (the crossing
  abscissa uses the edge's slope, a scalar, exactly as the reference does);
* links: (fluid neighbour, direction towards the solid) pairs of every solid
  node, skipping neighbours beyond the upper lattice bounds and wrapping
  negative ones like NumPy indexing, then ``np.unique(axis=0)`` (lattice.py:328-344);
* IBB distance: distance from the fluid node to the nearest polygon sample,
  divided by dx*|c_q| (lattice.py:350-361).
"""
import numpy as np

C = np.array([[0, 0], [1, 0], [-1, 0], [0, 1], [0, -1], [1, 1], [-1, -1], [-1, 1], [1, -1]])
NS = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7])


def node_coords(lat, i, j):
    """Coordinates of nodes (i, j) (lattice.py:379-387); i, j may be arrays."""
    dx = (lat.x_max - lat.x_min) / (lat.nx - 1)
    dy = (lat.y_max - lat.y_min) / (lat.ny - 1)
    return lat.x_min + np.asarray(i, dtype=np.float64) * dx, \
        lat.y_min + np.asarray(j, dtype=np.float64) * dy


def inside_polygon(poly, px, py):
    """Even-odd crossing test of points (px, py) against a closed polygon."""
    odd = np.zeros(px.shape, dtype=bool)
    n = len(poly)
    for i in range(n):
        j = i - 1 if i > 0 else n - 1
        xi, yi, xj, yj = poly[i, 0], poly[i, 1], poly[j, 0], poly[j, 1]
        if yi == yj:
            continue  # the reference's straddle condition is never true
        straddle = ((yi < py) & (yj >= py)) | ((yj < py) & (yi >= py))
        left = (xi < px) | (xj < px)
        slope = (xj - xi) / (yj - yi)
        cross = (xi + (py - yi) * slope) < px
        odd ^= straddle & left & cross
    return odd


def rasterise(lat, polygon, tag):
    """Tag the nodes inside ``polygon`` in ``lat.lattice``; return them as (i, j)
    pairs in the reference's i-major order."""
    lo_x, hi_x = np.amin(polygon[:, 0]), np.amax(polygon[:, 0])
    lo_y, hi_y = np.amin(polygon[:, 1]), np.amax(polygon[:, 1])
    # index window of the bounding box, two nodes wider than needed; the exact
    # strict comparisons below decide membership as the reference does
    gx = (lat.x_max - lat.x_min) / (lat.nx - 1)
    gy = (lat.y_max - lat.y_min) / (lat.ny - 1)
    i0 = max(0, int(np.floor((lo_x - lat.x_min) / gx)) - 2)
    i1 = min(lat.nx, int(np.ceil((hi_x - lat.x_min) / gx)) + 3)
    j0 = max(0, int(np.floor((lo_y - lat.y_min) / gy)) - 2)
    j1 = min(lat.ny, int(np.ceil((hi_y - lat.y_min) / gy)) + 3)
    if i1 <= i0 or j1 <= j0:
        return np.empty((0, 2), dtype=int)
    ii, jj = np.meshgrid(np.arange(i0, i1), np.arange(j0, j1), indexing="ij")
    px, py = node_coords(lat, ii, jj)
    box = (px > lo_x) & (px < hi_x) & (py > lo_y) & (py < hi_y)
    bi, bj = ii[box], jj[box]
    inside = inside_polygon(polygon, px[box], py[box])
    oi, oj = bi[inside], bj[inside]
    lat.lattice[oi, oj] = tag
    return np.stack([oi, oj], axis=1).astype(int)


def boundary_links(lat, solid):
    """Sorted unique links (i, j, q), q pointing from the fluid node to the solid."""
    if len(solid) == 0:
        return np.empty((0, 3), dtype=int)
    i = solid[:, 0:1]
    j = solid[:, 1:2]
    ii = i + C[1:, 0][None, :]
    jj = j + C[1:, 1][None, :]
    qb = np.broadcast_to(NS[1:][None, :], ii.shape)
    ok = (ii <= lat.nx - 1) & (jj <= lat.ny - 1)
    ii, jj, qb = ii[ok], jj[ok], qb[ok]
    fluid = lat.lattice[ii, jj] == 0  # negative indices wrap as in NumPy
    bnd = np.stack([ii[fluid], jj[fluid], qb[fluid]], axis=1).astype(int)
    if len(bnd) == 0:
        return np.empty((0, 3), dtype=int)
    return np.unique(bnd, axis=0)


def ibb_distances(lat, polygon, bnd):
    """Nearest-sample distance / (dx |c_q|) of each link (lattice.py:350-361)."""
    if len(bnd) == 0:
        return np.empty((0), dtype=float)
    px, py = node_coords(lat, bnd[:, 0], bnd[:, 1])
    norm = np.sqrt(np.sum(C[bnd[:, 2]] ** 2, axis=1).astype(np.float64))
    out = np.empty(len(bnd))
    step = max(1, 4_000_000 // max(1, len(polygon)))
    for s in range(0, len(bnd), step):
        x = polygon[None, :, 0] - px[s:s + step, None]
        y = polygon[None, :, 1] - py[s:s + step, None]
        dist = np.sqrt(np.square(x) + np.square(y))
        k = np.argmin(dist, axis=1)
        out[s:s + step] = dist[np.arange(len(k)), k] / (lat.dx * norm[s:s + step])
    return out


def add_obstacle(lat, obstacle, verbose=True):
    """Rasterise an obstacle and build its links (lattice.py:290-375).
    Returns (area, boundary int[nb,3], ibb float[nb])."""
    tag = obstacle.tag
    if verbose:
        print('# Obstacle ', str(tag))
    polygon = obstacle.polygon
    solid = rasterise(lat, polygon, tag)
    if verbose:
        print('# ' + str(solid.shape[0]) + ' locations in obstacle')
    bnd = boundary_links(lat, solid)
    if verbose:
        print('# ' + str(bnd.shape[0]) + ' locations on boundary')
    ibb = ibb_distances(lat, polygon, bnd) if lat.IBB else np.empty((0), dtype=float)
    # the reference accumulates dx**2 node by node
    area = 0.0
    d2 = lat.dx ** 2
    for _ in range(int(np.count_nonzero(lat.lattice == tag))):
        area += d2
    if verbose:
        print('# Area = ' + '{:f}'.format(area))
        print('')
    return area, bnd, ibb
