"""PyTorch-eager FP64 restatement of the partial-bounce-back channel run --
TEST INFRASTRUCTURE (the checker of lbm_amd.adjoint).

Written on full populations with the reference's formulas (lattice.py:178-189,
nb.py:7-45, the Zou-He routines nb.py:121-344 in the channel apps' order
turek.py:112-125), plus partial bounce-back P = (1-s) g_coll + s g_opposite
and drag D = sum_x s sum_q c_qx (g_coll_q - g_opposite_q).  Differentiated by
torch.autograd; independent of the GPU code's deviation form and of its
hand-written adjoint.
"""
import torch

CX = [0, 1, -1, 0, 0, 1, -1, -1, 1]
CY = [0, 0, 0, 1, -1, 1, -1, 1, -1]
NS = [0, 2, 1, 4, 3, 6, 5, 8, 7]
W = [4. / 9.] + [1. / 9.] * 4 + [1. / 36.] * 4
K1, K2, K3 = 2.0 / 3.0, 1.0 / 6.0, 1.0 / 2.0


def _set(plane, idx, val):
    """plane with plane[idx] = val (functional, autograd-friendly)."""
    out = plane.clone()
    out[idx] = val
    return out


def _walls(f, nx, ny, ux_in, uy_in, rho_out):
    lx, ly = nx - 1, ny - 1
    f = list(f)
    # bottom (y = 0, u = 0)
    r = (f[0][:, 0] + f[1][:, 0] + f[2][:, 0] + 2 * f[4][:, 0] + 2 * f[6][:, 0] + 2 * f[8][:, 0])
    f3 = f[4][:, 0]
    f5 = f[6][:, 0] - K3 * (f[1][:, 0] - f[2][:, 0])
    f7 = f[8][:, 0] + K3 * (f[1][:, 0] - f[2][:, 0])
    f[3], f[5], f[7] = _set(f[3], (slice(None), 0), f3), _set(f[5], (slice(None), 0), f5), \
        _set(f[7], (slice(None), 0), f7)
    rho_b = r
    # left (x = 0, u = inlet)
    r = (f[0][0] + f[3][0] + f[4][0] + 2 * f[2][0] + 2 * f[6][0] + 2 * f[7][0]) / (1 - ux_in)
    f1 = f[2][0] + K1 * r * ux_in
    f5 = f[6][0] - K3 * (f[3][0] - f[4][0]) + K2 * r * ux_in + K3 * r * uy_in
    f8 = f[7][0] + K3 * (f[3][0] - f[4][0]) + K2 * r * ux_in - K3 * r * uy_in
    f[1], f[5], f[8] = _set(f[1], 0, f1), _set(f[5], 0, f5), _set(f[8], 0, f8)
    # top (y = ly, u = 0)
    r = (f[0][:, ly] + f[1][:, ly] + f[2][:, ly] + 2 * f[3][:, ly] + 2 * f[5][:, ly]
         + 2 * f[7][:, ly])
    f4 = f[3][:, ly]
    f8 = f[7][:, ly] - K3 * (f[1][:, ly] - f[2][:, ly])
    f6 = f[5][:, ly] + K3 * (f[1][:, ly] - f[2][:, ly])
    f[4], f[8], f[6] = _set(f[4], (slice(None), ly), f4), _set(f[8], (slice(None), ly), f8), \
        _set(f[6], (slice(None), ly), f6)
    rho_t = r
    # right (x = lx, pressure)
    ux = (f[0][lx] + f[3][lx] + f[4][lx] + 2 * f[1][lx] + 2 * f[5][lx] + 2 * f[8][lx]) / rho_out - 1
    f2 = f[1][lx] - K1 * rho_out * ux
    f6 = f[5][lx] + K3 * (f[3][lx] - f[4][lx]) - K2 * rho_out * ux
    f7 = f[8][lx] - K3 * (f[3][lx] - f[4][lx]) - K2 * rho_out * ux
    f[2], f[6], f[7] = _set(f[2], lx, f2), _set(f[6], lx, f6), _set(f[7], lx, f7)
    # corners: rho of the adjacent bottom/top wall node, u = 0
    for (i, j, r, pairs, zeros) in (
            (0, 0, rho_b[1], ((1, 2), (3, 4), (5, 6)), (7, 8)),
            (0, ly, rho_t[1], ((1, 2), (4, 3), (8, 7)), (5, 6)),
            (lx, ly, rho_t[lx - 1], ((2, 1), (4, 3), (6, 5)), (7, 8)),
            (lx, 0, rho_b[lx - 1], ((2, 1), (3, 4), (7, 8)), (5, 6))):
        for a, b in pairs:
            f[a] = _set(f[a], (i, j), f[b][i, j])
        for z in zeros:
            f[z] = _set(f[z], (i, j), torch.zeros((), dtype=f[z].dtype))
        rest = r
        for q in range(1, 9):
            rest = rest - f[q][i, j]
        f[0] = _set(f[0], (i, j), rest)
    return f


def _collide(f, om_p, om_m):
    rho = sum(f)
    ux = sum(CX[q] * f[q] for q in range(9)) / rho
    uy = sum(CY[q] * f[q] for q in range(9)) / rho
    v = 1.5 * (ux ** 2 + uy ** 2)
    geq = []
    for q in range(9):
        t = 3.0 * (ux * CX[q] + uy * CY[q])
        geq.append((1.0 + t + 0.5 * t ** 2 - v) * rho * W[q])
    out = [(1.0 - om_p) * f[0] + om_p * geq[0]]
    for q in range(1, 9):
        qb = NS[q]
        out.append((1.0 - 0.5 * (om_p + om_m)) * f[q] - 0.5 * (om_p - om_m) * f[qb]
                   + 0.5 * (om_p + om_m) * geq[q] + 0.5 * (om_p - om_m) * geq[qb])
    return out


def run(case, s, u_max, iterations=None):
    """FP64 forward on the CPU; returns (loss, per-iteration drag, final P - w)."""
    nx, ny = case["nx"], case["ny"]
    T = case["iterations"] if iterations is None else iterations
    om_p, om_m = case["om_p"], case["om_m"]
    prof = torch.as_tensor(case["profile"], dtype=torch.float64)
    ramp = case["ramp"]
    f = [torch.full((nx, ny), W[q], dtype=torch.float64) for q in range(9)]
    drags = []

    def pbb(f):
        c = _collide(f, om_p, om_m)
        P = [(1 - s) * c[q] + s * f[NS[q]] for q in range(9)]
        D = (s * sum(CX[q] * (c[q] - f[NS[q]]) for q in range(9))).sum()
        return P, D

    P, D = pbb(f)
    drags.append(D)
    for t in range(1, T + 1):
        fp = []
        for q in range(9):
            pad = torch.nn.functional.pad(P[q][None, None], (1, 1, 1, 1))[0, 0]
            fp.append(pad[1 - CX[q]:1 - CX[q] + nx, 1 - CY[q]:1 - CY[q] + ny])
        a = ramp[t - 1] * u_max
        f = _walls(fp, nx, ny, a * prof[0], a * prof[1], case.get("rho_out", 1.0))
        P, D = pbb(f)
        drags.append(D)
    window = case["window"]
    loss = sum(drags[T - window + 1:T + 1]) / window
    final = torch.stack([P[q] - W[q] for q in range(9)])
    return loss, torch.stack(drags), final
