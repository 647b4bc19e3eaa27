// d2q9_adjoint.hip -- channel step with a partial bounce-back obstacle
// (per-node solid fraction s in [0,1]) and its hand-written discrete adjoint.
//
// Forward iteration t (t >= 1), node x, in deviations h = f - w:
//   h   = pull(H_{t-1}) with the Zou-He plan of the channel apps
//         (turek.py:112-125: velocity walls bottom/top at rest, velocity
//         inlet u = a_{t-1} * profile on the left, pressure outlet on the
//         right, the four corners; nb.py:121-344)
//   c   = TRT collision of h (nb.py:7-35, written on deviations)
//   H_t = (1 - s) c + s h_bar          (partial bounce-back: h_bar_q = h_{ns[q]})
//   D_t = sum_x s sum_q c_qx (c_q - h_bar_q)   (momentum taken from the fluid)
// Iteration 0 starts from rest (h = 0, so H_0 = 0 for every s).
// Loss L = mean of D_t over the last `window` iterations.
//
// Adjoint iteration t, given lam_t = dL/dH_t, recomputes h, c at every node
// from H_{t-1} and forms
//   cbar = (1-s) lam + W s cx,   hbar = s lam_bar + W s cx,
//   sbar += lam . (h_bar - c) + W cx . (c - h_bar),
//   hbar += cbar - Om cbar + rhobar + c . jbar     (TRT + equilibrium VJP),
// then pushes hbar back along -c_q (the transpose of the pull); boundary
// nodes first apply the transpose of their (affine) Zou-He map, columns
// evaluated exactly, and the inlet nodes accumulate dL/da.
#include "d2q9_common.h"
#include "d2q9_zouhe.h"
#include "../../include/lbm_d2q9.h"

#include <memory>
#include <vector>

namespace lbm {
namespace {

struct PbbParams {
    int nx, ny, N;
    uint32_t magic;
    double om_p, om_m, rho_out;
    const double *prof;     // [2][ny] inlet profile of unit amplitude
    double amp;             // inlet amplitude a_{t-1} of this iteration's boundary step
    const void *src;        // H_{t-1}
    void *dst;              // H_t (forward)
    const void *s;          // solid fraction [N]
    double *drag_part;      // [gridDim.x] per-block drag sums (forward)
    const void *lam_in;     // lam_t (adjoint)
    void *lam_out;          // lam_{t-1} (adjoint)
    double *s_bar;          // [N] (adjoint)
    double *a_bar;          // [ny] (adjoint)
    double W;               // drag weight of this iteration
    double ramp;            // da/du_max of this iteration's inlet
    int init;               // forward iteration 0
};

template <typename T> __device__ __forceinline__ T rcp(T v) { return T(1) / v; }
template <> __device__ __forceinline__ float rcp<float>(float v) { return __builtin_amdgcn_rcpf(v); }

__device__ __forceinline__ int nodex(const PbbParams &p, int cell)
{
    int x = (int)__umulhi((uint32_t)cell, p.magic);
    if (cell - x * p.ny < 0) --x;
    return x;
}

template <typename T> __device__ __forceinline__ T ld(const void *b, size_t i)
{
    return static_cast<const T *>(b)[i];
}

// Boundary step of a non-corner perimeter node on full populations f
// (unknown entries are overwritten).
template <typename T>
__device__ __forceinline__ void edge_bc(const PbbParams &p, int x, int y, T *f)
{
    T r;
    const int lx = p.nx - 1, ly = p.ny - 1;
    if (y == 0) zh_bottom(f, T(0), T(0), r);
    else if (y == ly) zh_top(f, T(0), T(0), r);
    else if (x == 0) zh_left(f, T(p.amp * p.prof[y]), T(p.amp * p.prof[p.ny + y]), r);
    else if (x == lx) {
        T ux;
        zh_right_pres(f, T(p.rho_out), T(0), ux);
    }
}

// Density the bottom/top wall routine gives node (x, y) from its populations.
template <typename T>
__device__ __forceinline__ T wall_rho(int y, int ly, const T *fn)
{
    T g[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) g[q] = fn[q];
    T r;
    if (y == 0) zh_bottom(g, T(0), T(0), r);
    else zh_top(g, T(0), T(0), r);
    (void)ly;
    return r;
}

__device__ __forceinline__ int corner_of(int x, int y, int lx, int ly)
{
    if (x == 0) return y == 0 ? 0 : 1;
    return y == 0 ? 3 : 2;
}

__device__ __forceinline__ bool in_domain(const PbbParams &p, int x, int y)
{
    return x >= 0 && x < p.nx && y >= 0 && y < p.ny;
}

// Pulled full populations of node (x, y); entries without a source are w.
template <typename T>
__device__ __forceinline__ void pull_full(const PbbParams &p, const void *src, int x, int y, T *f)
{
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int sx = x - CX[q], sy = y - CY[q];
        f[q] = weight<T>(q);
        if (in_domain(p, sx, sy))
            f[q] += ld<T>(src, (size_t)q * p.N + (size_t)sx * p.ny + sy);
    }
}

// Post-boundary deviations h of any node.
template <typename T>
__device__ void node_h(const PbbParams &p, int x, int y, T *h)
{
    const int lx = p.nx - 1, ly = p.ny - 1;
    const bool interior = x > 0 && x < lx && y > 0 && y < ly;
    if (interior) {
#pragma unroll
        for (int q = 0; q < 9; ++q)
            h[q] = ld<T>(p.src, (size_t)q * p.N + (size_t)(x - CX[q]) * p.ny + (y - CY[q]));
        return;
    }
    T f[9];
    pull_full<T>(p, p.src, x, y, f);
    const bool corner = (x == 0 || x == lx) && (y == 0 || y == ly);
    if (!corner) {
        edge_bc<T>(p, x, y, f);
    } else {
        const int xn = x == 0 ? 1 : lx - 1;
        T fn[9];
        pull_full<T>(p, p.src, xn, y, fn);
        zh_corner(f, corner_of(x, y, lx, ly), wall_rho<T>(y, ly, fn), T(0), T(0));
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) h[q] = f[q] - weight<T>(q);
}

// TRT collision on deviations; also returns rho and u.
template <typename T>
__device__ __forceinline__ void collide_dev(T om_p, T om_m, const T *h, T *c, T &rho, T &ux,
                                            T &uy)
{
    const T drho = ((h[0] + h[1]) + (h[2] + h[3])) + ((h[4] + h[5]) + (h[6] + h[7])) + h[8];
    const T jx = (h[1] - h[2]) + (h[5] - h[6]) + (h[8] - h[7]);
    const T jy = (h[3] - h[4]) + (h[5] - h[6]) + (h[7] - h[8]);
    rho = T(1) + drho;
    const T inv = rcp(rho);
    ux = jx * inv;
    uy = jy * inv;
    const T usq = T(1.5) * (ux * ux + uy * uy);
    const T base = drho - rho * usq;
    c[0] = h[0] - om_p * (h[0] - weight<T>(0) * base);
    const T hp = T(0.5) * om_p, hm = T(0.5) * om_m;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int q = 2 * j + 1, qb = q + 1;
        const T w = weight<T>(q);
        const T cu = j == 0 ? ux : (j == 1 ? uy : (j == 2 ? ux + uy : uy - ux));
        const T ep = w * (base + T(4.5) * rho * cu * cu);
        const T em = T(3) * w * rho * cu;
        const T dp = hp * ((h[q] + h[qb]) - T(2) * ep);
        const T dm = hm * ((h[q] - h[qb]) - T(2) * em);
        c[q] = h[q] - dp - dm;
        c[qb] = h[qb] - dp + dm;
    }
}

// Adjoint of the collision + partial bounce-back + drag at one node:
// lam (adjoint of H at the node) -> hbar (adjoint of the post-boundary h),
// and the node's contribution to dL/ds.
template <typename T>
__device__ __forceinline__ void node_adjoint(const PbbParams &p, const T *h, T sv, const T *lam,
                                             T *hbar, double &sbar)
{
    const T om_p = T(p.om_p), om_m = T(p.om_m), W = T(p.W);
    T c[9], rho, ux, uy;
    collide_dev<T>(om_p, om_m, h, c, rho, ux, uy);
    T cb[9];
    double sb = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int qb = NS[q];
        cb[q] = (T(1) - sv) * lam[q] + W * sv * T(CX[q]);
        hbar[q] = sv * lam[qb] + W * sv * T(CX[q]);
        sb += (double)(lam[q] * (h[qb] - c[q]) + W * T(CX[q]) * (c[q] - h[qb]));
    }
    sbar = sb;
    // Ebar = Om cbar (TRT relaxation operator, symmetric)
    T eb[9];
    eb[0] = om_p * cb[0];
#pragma unroll
    for (int q = 1; q < 9; q += 2) {
        const T sp = T(0.5) * (cb[q] + cb[q + 1]), sm = T(0.5) * (cb[q] - cb[q + 1]);
        eb[q] = om_p * sp + om_m * sm;
        eb[q + 1] = om_p * sp - om_m * sm;
    }
    // E_q = w (drho + 3 c.j + (4.5 (c.j)^2 - 1.5 j.j) / rho)
    const T u2 = ux * ux + uy * uy;
    T rb = T(0), jbx = T(0), jby = T(0);
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const T w = weight<T>(q);
        const T cu = T(CX[q]) * ux + T(CY[q]) * uy;
        rb += eb[q] * w * (T(1) - T(4.5) * cu * cu + T(1.5) * u2);
        const T k = eb[q] * w;
        jbx += k * (T(3) * T(CX[q]) + T(9) * cu * T(CX[q]) - T(3) * ux);
        jby += k * (T(3) * T(CY[q]) + T(9) * cu * T(CY[q]) - T(3) * uy);
    }
#pragma unroll
    for (int q = 0; q < 9; ++q)
        hbar[q] += cb[q] - eb[q] + rb + T(CX[q]) * jbx + T(CY[q]) * jby;
}

// ------------------------------------------------------------- forward --
template <typename T>
__global__ __launch_bounds__(256) void k_pbb_fwd(PbbParams p)
{
    __shared__ double red[256];
    const int cell = blockIdx.x * 256 + threadIdx.x;
    double d = 0.0;
    if (cell < p.N) {
        const int x = nodex(p, cell), y = cell - x * p.ny;
        T h[9];
        if (p.init) {
#pragma unroll
            for (int q = 0; q < 9; ++q) h[q] = T(0);
        } else {
            node_h<T>(p, x, y, h);
        }
        T c[9], rho, ux, uy;
        collide_dev<T>(T(p.om_p), T(p.om_m), h, c, rho, ux, uy);
        const T sv = ld<T>(p.s, cell);
        T* dst = static_cast<T *>(p.dst);
        double dd = 0.0;
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const int qb = NS[q];
            dst[(size_t)q * p.N + cell] = (T(1) - sv) * c[q] + sv * h[qb];
            dd += (double)(T(CX[q]) * (c[q] - h[qb]));
        }
        d = (double)sv * dd;
    }
    red[threadIdx.x] = d;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0 && p.drag_part) p.drag_part[blockIdx.x] = red[0];
}

// ------------------------------------------------------------- adjoint --
// Transposed boundary step of a perimeter node: hbar (adjoint of its
// post-boundary deviations) -> adjoint of its pulled populations (own) and,
// for a corner, of the adjacent wall node's pulled populations.  The boundary
// map is affine in the pulled populations, so its columns are exact
// differences of two evaluations.
// Column k of that transposed map: sum_q hbar_q d(out_q)/d(in_k), with `in`
// the node's own pulled populations or, for a corner (nbr), those of the
// adjacent wall node.
template <typename T>
__device__ T bc_column(const PbbParams &p, int x, int y, const T *hbar, int k, bool nbr)
{
    const int lx = p.nx - 1, ly = p.ny - 1;
    const bool corner = (x == 0 || x == lx) && (y == 0 || y == ly);
    const int which = corner ? corner_of(x, y, lx, ly) : 0;
    auto eval = [&](const T *fo, const T *fnb, T *out) {
#pragma unroll
        for (int q = 0; q < 9; ++q) out[q] = fo[q];
        if (!corner) {
            edge_bc<T>(p, x, y, out);
        } else {
            zh_corner(out, which, wall_rho<T>(y, ly, fnb), T(0), T(0));
        }
    };
    const int xs = nbr ? (x == 0 ? 1 : lx - 1) : x;
    if (!in_domain(p, xs - CX[k], y - CY[k])) return T(0);   // not pulled: no adjoint
    T zero[9], e[9], base[9], col[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        zero[q] = T(0);
        e[q] = q == k ? T(1) : T(0);
    }
    eval(zero, zero, base);
    if (nbr) eval(zero, e, col);
    else eval(e, zero, col);
    T acc = T(0);
#pragma unroll
    for (int q = 0; q < 9; ++q) acc += hbar[q] * (col[q] - base[q]);
    return acc;
}

// d(boundary output)/d(ux, uy) . hbar on a left-wall node (zh_left with
// rho = S / (1 - ux)).
template <typename T>
__device__ __forceinline__ void left_velocity_adjoint(const T *f, T ux, T uy, const T *hbar,
                                                      T &gx, T &gy)
{
    const T S = f[0] + f[3] + f[4] + T(2) * f[2] + T(2) * f[6] + T(2) * f[7];
    const T r = S / (T(1) - ux);
    const T dr = r / (T(1) - ux);          // d(rho ux)/dux = rho / (1 - ux)
    const T d1x = T(K1) * dr;
    const T d5x = T(K2) * dr + T(K3) * uy * (r / (T(1) - ux));
    const T d8x = T(K2) * dr - T(K3) * uy * (r / (T(1) - ux));
    gx = hbar[1] * d1x + hbar[5] * d5x + hbar[8] * d8x;
    gy = hbar[5] * (T(K3) * r) - hbar[8] * (T(K3) * r);
}

// Interior node (every pull in the domain, no boundary map).
template <typename T>
__device__ __forceinline__ void adj_interior_node(const PbbParams &p, int cell)
{
    if (cell >= p.N) return;
    const int x = nodex(p, cell), y = cell - x * p.ny;
    if (x == 0 || x == p.nx - 1 || y == 0 || y == p.ny - 1) return;
    T h[9], lam[9], hb[9];
#pragma unroll
    for (int q = 0; q < 9; ++q)
        h[q] = ld<T>(p.src, (size_t)q * p.N + (size_t)(x - CX[q]) * p.ny + (y - CY[q]));
#pragma unroll
    for (int q = 0; q < 9; ++q) lam[q] = ld<T>(p.lam_in, (size_t)q * p.N + cell);
    const T sv = ld<T>(p.s, cell);
    double sb;
    node_adjoint<T>(p, h, sv, lam, hb, sb);
    p.s_bar[cell] += sb;
    T *lo = static_cast<T *>(p.lam_out);
#pragma unroll
    for (int q = 0; q < 9; ++q)
        lo[(size_t)q * p.N + (size_t)(x - CX[q]) * p.ny + (y - CY[q])] = hb[q];
}

// Perimeter node i (bottom row, top row, left and right columns without the
// corners) -> (x, y).
__device__ __forceinline__ void pbb_perim_node(const PbbParams &p, int i, int &x, int &y)
{
    if (i < p.nx) { x = i; y = 0; return; }
    i -= p.nx;
    if (i < p.nx) { x = i; y = p.ny - 1; return; }
    i -= p.nx;
    if (i < p.ny - 2) { x = 0; y = i + 1; return; }
    i -= p.ny - 2;
    x = p.nx - 1;
    y = i + 1;
}

// Perimeter nodes: boundary transposes, corner terms, inlet sensitivity.
// One work-item per (perimeter node, plane k): each recomputes the node's
// adjoint and evaluates one column of its boundary map, so the nine columns
// run in parallel instead of in one long dependent chain per node.
template <typename T>
__device__ __forceinline__ void adj_perimeter_item(const PbbParams &p, int i)
{
    const int np = 2 * p.nx + 2 * p.ny - 4;
    if (i >= 9 * np) return;
    const int node = i / 9, k = i - 9 * node;
    int x, y;
    pbb_perim_node(p, node, x, y);
    const int cell = x * p.ny + y;
    const int lx = p.nx - 1, ly = p.ny - 1;
    T h[9], lam[9], hb[9];
    node_h<T>(p, x, y, h);
#pragma unroll
    for (int q = 0; q < 9; ++q) lam[q] = ld<T>(p.lam_in, (size_t)q * p.N + cell);
    double sb;
    node_adjoint<T>(p, h, ld<T>(p.s, cell), lam, hb, sb);
    T out = bc_column<T>(p, x, y, hb, k, false);
    if (k == 0) {
        p.s_bar[cell] += sb;
        if (x == 0 && y > 0 && y < ly) {
            T f[9];
            pull_full<T>(p, p.src, x, y, f);
            T gx, gy;
            left_velocity_adjoint<T>(f, T(p.amp * p.prof[y]), T(p.amp * p.prof[p.ny + y]), hb,
                                     gx, gy);
            p.a_bar[y] += p.ramp * ((double)gx * p.prof[y] + (double)gy * p.prof[p.ny + y]);
        }
    }
    // a wall node next to a corner also carries the corner's dependence on it
    if ((y == 0 || y == ly) && (x == 1 || x == lx - 1)) {
        const int xc = x == 1 ? 0 : lx;
        const int cc = xc * p.ny + y;
        T hc[9], lc[9], hbc[9];
        node_h<T>(p, xc, y, hc);
#pragma unroll
        for (int q = 0; q < 9; ++q) lc[q] = ld<T>(p.lam_in, (size_t)q * p.N + cc);
        double sdummy;
        node_adjoint<T>(p, hc, ld<T>(p.s, cc), lc, hbc, sdummy);
        out += bc_column<T>(p, xc, y, hbc, k, true);
    }
    T *lo = static_cast<T *>(p.lam_out);
    const int sx = x - CX[k], sy = y - CY[k];
    if (in_domain(p, sx, sy)) lo[(size_t)k * p.N + (size_t)sx * p.ny + sy] = out;
    // an entry of this node that no node pulls carries no adjoint
    if (!in_domain(p, x + CX[k], y + CY[k])) lo[(size_t)k * p.N + cell] = T(0);
}

// One adjoint iteration in one launch: work-groups [0, eb) take the
// perimeter items (a long dependent chain each, so they start first and run
// beside the interior), the rest one interior node per work-item.  The two
// parts write disjoint entries: lam_out[q][m] is written by the one node that
// pulls it, s_bar per node, a_bar by perimeter nodes only.  W is an occupancy
// floor that keeps the interior path's register count (FP64: 4 waves/SIMD,
// 128 VGPRs; the perimeter path alone would need 158 and spills a little).
template <typename T, int W>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W))) void k_pbb_adj(
    PbbParams p, int eb)
{
    // interior node blocks dealt to the XCDs four at a time (neighbouring
    // rows in one L2; FP64 backward 64.2 -> 63.0 ms on the 1024x256 channel)
    if ((int)blockIdx.x < eb)
        adj_perimeter_item<T>(p, blockIdx.x * 256 + threadIdx.x);
    else
        adj_interior_node<T>(
            p, xcd_block_unit((int)blockIdx.x - eb, (int)gridDim.x - eb, 2) * 256 + threadIdx.x);
}

__global__ void k_pbb_sum_rows(const double *part, int nb, int rows, double *out)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    double s = 0.0;
    for (int i = 0; i < nb; ++i) s += part[(size_t)r * nb + i];
    out[r] = s;
}

__global__ void k_zero_f64(double *p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

template <typename T> __global__ void k_zero_t(T *p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = T(0);
}

inline hipStream_t SS(void *s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace
}  // namespace lbm

using namespace lbm;

struct lbm_pbb {
    int nx = 0, ny = 0, N = 0, storage = 0, T = 0, window = 0, every = 0, nblocks = 0;
    uint32_t magic = 0;
    double om_p = 0, om_m = 0, rho_out = 1;
    double *d_prof = nullptr;
    std::vector<double> ramp;          // ramp(it), it = 0..T
    double u_max = 0;
    size_t sb = 0;                     // bytes of one state
    std::vector<void *> ckpt;          // H at it = 0, every, 2 every, ...
    void *ping[2] = {nullptr, nullptr};
    void *stack = nullptr;             // one segment of states
    void *lam[2] = {nullptr, nullptr};
    double *d_part = nullptr;          // [T+1][nblocks]
    double *d_drag = nullptr;          // [T+1]
    double *d_abar = nullptr;          // [ny]
    const void *s = nullptr;           // solid fraction of the last forward
    std::vector<double> drag;          // host copy
};

namespace {

PbbParams pbb_params(const lbm_pbb *h)
{
    PbbParams p{};
    p.nx = h->nx;
    p.ny = h->ny;
    p.N = h->N;
    p.magic = h->magic;
    p.om_p = h->om_p;
    p.om_m = h->om_m;
    p.rho_out = h->rho_out;
    p.prof = h->d_prof;
    p.s = h->s;
    return p;
}

void pbb_free(lbm_pbb *h)
{
    for (void *c : h->ckpt) (void)hipFree(c);
    void *ptrs[] = {h->d_prof, h->ping[0], h->ping[1], h->stack, h->lam[0], h->lam[1],
                    h->d_part, h->d_drag, h->d_abar};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
}

template <typename T> void launch_fwd(const lbm_pbb *h, const PbbParams &p, hipStream_t s)
{
    k_pbb_fwd<T><<<h->nblocks, 256, 0, s>>>(p);
}

void fwd(const lbm_pbb *h, const PbbParams &p, hipStream_t s)
{
    if (h->storage == LBM_STORE_F64) launch_fwd<double>(h, p, s);
    else launch_fwd<float>(h, p, s);
}

void adj(const lbm_pbb *h, const PbbParams &p, hipStream_t s)
{
    const int eb = (9 * (2 * h->nx + 2 * h->ny - 4) + 255) / 256;
    if (h->storage == LBM_STORE_F64)
        k_pbb_adj<double, 4><<<eb + h->nblocks, 256, 0, s>>>(p, eb);
    else
        k_pbb_adj<float, 8><<<eb + h->nblocks, 256, 0, s>>>(p, eb);
}

// state buffer holding H_t during the forward sweep
void *state_of(lbm_pbb *h, int t)
{
    if (t % h->every == 0) return h->ckpt[t / h->every];
    return h->ping[t & 1];
}

}  // namespace

extern "C" int lbm_pbb_create(const lbm_pbb_desc *d, lbm_pbb **out)
{
    if (!d || !out || d->nx < 4 || d->ny < 4 || d->iterations < 1 || d->window < 1 ||
        d->window > d->iterations || d->checkpoint_every < 1 || !d->inlet_profile_host ||
        !d->ramp_host || (d->storage != LBM_STORE_F64 && d->storage != LBM_STORE_F32))
        return set_error(LBM_ERR_ARG, "lbm_pbb_create: bad argument");
    if ((long long)d->nx * d->ny * 9 >= (1LL << 31))
        return set_error(LBM_ERR_ARG, "lattice too large");
    auto h = std::make_unique<lbm_pbb>();
    h->nx = d->nx;
    h->ny = d->ny;
    h->N = d->nx * d->ny;
    h->storage = d->storage;
    h->T = d->iterations;
    h->window = d->window;
    h->every = d->checkpoint_every;
    h->om_p = d->om_p;
    h->om_m = d->om_m;
    h->rho_out = d->rho_out;
    h->magic = (uint32_t)((1ULL << 32) / (uint64_t)d->ny) + 1u;
    h->nblocks = (h->N + 255) / 256;
    h->ramp.assign(d->ramp_host, d->ramp_host + d->iterations + 1);
    h->sb = (size_t)9 * h->N * (d->storage == LBM_STORE_F64 ? 8 : 4);
    auto alloc = [&](void **p, size_t bytes) {
        return hipMalloc(p, bytes) == hipSuccess;
    };
    const int nck = h->T / h->every + 1;
    h->ckpt.assign(nck, nullptr);
    bool ok = alloc((void **)&h->d_prof, sizeof(double) * 2 * h->ny) &&
              alloc(&h->ping[0], h->sb) && alloc(&h->ping[1], h->sb) &&
              alloc(&h->stack, h->sb * (size_t)h->every) && alloc(&h->lam[0], h->sb) &&
              alloc(&h->lam[1], h->sb) &&
              alloc((void **)&h->d_part, sizeof(double) * (size_t)(h->T + 1) * h->nblocks) &&
              alloc((void **)&h->d_drag, sizeof(double) * (h->T + 1)) &&
              alloc((void **)&h->d_abar, sizeof(double) * h->ny);
    for (int i = 0; ok && i < nck; ++i) ok = alloc(&h->ckpt[i], h->sb);
    if (!ok) {
        pbb_free(h.get());
        return set_error(LBM_ERR_HIP, "lbm_pbb_create: allocation failed");
    }
    if (hipMemcpy(h->d_prof, d->inlet_profile_host, sizeof(double) * 2 * h->ny,
                  hipMemcpyHostToDevice) != hipSuccess) {
        pbb_free(h.get());
        return set_error(LBM_ERR_HIP, "lbm_pbb_create: profile upload");
    }
    *out = h.release();
    return LBM_OK;
}

extern "C" void lbm_pbb_destroy(lbm_pbb *h)
{
    if (!h) return;
    (void)hipDeviceSynchronize();
    pbb_free(h);
    delete h;
}

extern "C" int lbm_pbb_forward(lbm_pbb *h, const void *s_dev, double u_max, double *loss,
                               void *stream)
{
    if (!h || !s_dev || !loss) return set_error(LBM_ERR_ARG, "lbm_pbb_forward: bad argument");
    hipStream_t st = SS(stream);
    h->s = s_dev;
    h->u_max = u_max;
    PbbParams p = pbb_params(h);
    // iteration 0: from rest
    p.init = 1;
    p.dst = state_of(h, 0);
    p.drag_part = h->d_part;
    fwd(h, p, st);
    LBM_LAUNCH_CHECK();
    p.init = 0;
    for (int t = 1; t <= h->T; ++t) {
        p.src = state_of(h, t - 1);
        p.dst = state_of(h, t);
        p.amp = h->ramp[t - 1] * u_max;
        p.drag_part = h->d_part + (size_t)t * h->nblocks;
        fwd(h, p, st);
    }
    LBM_LAUNCH_CHECK();
    k_pbb_sum_rows<<<(h->T + 1 + 127) / 128, 128, 0, st>>>(h->d_part, h->nblocks, h->T + 1,
                                                          h->d_drag);
    LBM_LAUNCH_CHECK();
    h->drag.assign(h->T + 1, 0.0);
    LBM_HIP_TRY(hipMemcpyAsync(h->drag.data(), h->d_drag, sizeof(double) * (h->T + 1),
                               hipMemcpyDeviceToHost, st));
    LBM_HIP_TRY(hipStreamSynchronize(st));
    double acc = 0.0;
    for (int t = h->T - h->window + 1; t <= h->T; ++t) acc += h->drag[t];
    *loss = acc / h->window;
    return LBM_OK;
}

extern "C" int lbm_pbb_backward(lbm_pbb *h, double *s_grad_dev, double *umax_grad, void *stream)
{
    if (!h || !s_grad_dev || !umax_grad || !h->s)
        return set_error(LBM_ERR_ARG, "lbm_pbb_backward: call lbm_pbb_forward first");
    hipStream_t st = SS(stream);
    const size_t nstate = (size_t)9 * h->N;
    const int zb = (int)((nstate + 255) / 256);
    if (h->storage == LBM_STORE_F64)
        k_zero_t<double><<<zb, 256, 0, st>>>(static_cast<double *>(h->lam[h->T & 1]), nstate);
    else
        k_zero_t<float><<<zb, 256, 0, st>>>(static_cast<float *>(h->lam[h->T & 1]), nstate);
    k_zero_f64<<<(h->N + 255) / 256, 256, 0, st>>>(s_grad_dev, h->N);
    k_zero_f64<<<(h->ny + 255) / 256, 256, 0, st>>>(h->d_abar, h->ny);
    LBM_LAUNCH_CHECK();
    PbbParams base = pbb_params(h);
    base.s_bar = s_grad_dev;
    base.a_bar = h->d_abar;
    const double wgt = 1.0 / h->window;
    char *stack = static_cast<char *>(h->stack);
    // segments [t0, t1] of states, last first
    for (int t0 = ((h->T - 1) / h->every) * h->every; t0 >= 0; t0 -= h->every) {
        const int t1 = std::min(t0 + h->every - 1, h->T - 1);
        // recompute H_{t0+1} .. H_{t1} from the checkpoint H_{t0} into the stack
        {
            PbbParams p = base;
            p.init = 0;
            p.drag_part = nullptr;
            for (int t = t0 + 1; t <= t1; ++t) {
                p.src = (t - 1 == t0) ? h->ckpt[t0 / h->every]
                                      : (void *)(stack + h->sb * (size_t)(t - 1 - t0));
                p.dst = stack + h->sb * (size_t)(t - t0);
                p.amp = h->ramp[t - 1] * h->u_max;
                fwd(h, p, st);
            }
            LBM_LAUNCH_CHECK();
        }
        // adjoint iterations t = t1+1 .. t0+1, reading H_{t-1}
        for (int t = t1 + 1; t >= t0 + 1; --t) {
            PbbParams p = base;
            p.src = (t - 1 == t0) ? h->ckpt[t0 / h->every]
                                  : (void *)(stack + h->sb * (size_t)(t - 1 - t0));
            p.amp = h->ramp[t - 1] * h->u_max;
            p.ramp = h->ramp[t - 1];
            p.W = (t > h->T - h->window) ? wgt : 0.0;
            p.lam_in = h->lam[t & 1];
            p.lam_out = h->lam[(t - 1) & 1];
            adj(h, p, st);
        }
        LBM_LAUNCH_CHECK();
    }
    std::vector<double> ab(h->ny);
    LBM_HIP_TRY(hipMemcpyAsync(ab.data(), h->d_abar, sizeof(double) * h->ny,
                               hipMemcpyDeviceToHost, st));
    LBM_HIP_TRY(hipStreamSynchronize(st));
    double g = 0.0;
    for (int j = 0; j < h->ny; ++j) g += ab[j];
    *umax_grad = g;
    return LBM_OK;
}

extern "C" int lbm_pbb_drag_history(const lbm_pbb *h, double *out_host)
{
    if (!h || !out_host || h->drag.empty())
        return set_error(LBM_ERR_ARG, "lbm_pbb_drag_history: no forward run");
    std::copy(h->drag.begin(), h->drag.end(), out_host);
    return LBM_OK;
}

extern "C" int lbm_pbb_final_state(lbm_pbb *h, void *dst_dev, void *stream)
{
    if (!h || !dst_dev) return set_error(LBM_ERR_ARG, "lbm_pbb_final_state: bad argument");
    LBM_HIP_TRY(hipMemcpyAsync(dst_dev, state_of(h, h->T), h->sb, hipMemcpyDeviceToDevice,
                               SS(stream)));
    return LBM_OK;
}
