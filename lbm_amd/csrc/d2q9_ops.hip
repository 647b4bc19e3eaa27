// d2q9_ops.hip -- operation-by-operation D2Q9 kernels (double precision).
//
// One kernel (or kernel pair) per reference function of lbm/src/core/nb.py
// and lattice.macro, with the same in-place semantics, so that a caller that
// drives the lattice call by call (lattice.py:178-286) gets the reference's
// state after every call.  Expressions follow the reference's evaluation
// order; the library is compiled with -ffp-contract=off, so results are
// bit-identical to a strict IEEE evaluation of the reference expressions.
#include "d2q9_common.h"
#include "../../include/lbm_d2q9.h"

namespace lbm {
namespace {

constexpr int BLOCK = 256;

inline int nblocks(long long n) { return (int)((n + BLOCK - 1) / BLOCK); }

// lattice.py:178-189.  rho = sum_q g_q in q order (np.sum over axis 0);
// u = sum_q c_q g_q / rho, the BLAS dot products of np.tensordot, whose
// products are exact (c in {-1,0,1}) and whose sum runs in q order.
__global__ void k_macro(int nx, int ny, const double *__restrict__ g,
                        double *__restrict__ rho, double *__restrict__ u)
{
    const long long N = (long long)nx * ny;
    long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= N) return;
    double f[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) f[q] = g[q * N + i];
    double r = f[0];
#pragma unroll
    for (int q = 1; q < 9; ++q) r = r + f[q];
    double sx = 0.0, sy = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        sx = sx + (double)CX[q] * f[q];
        sy = sy + (double)CY[q] * f[q];
    }
    rho[i] = r;
    u[i] = sx / r;
    u[N + i] = sy / r;
}

// nb.py:7-17
__global__ void k_equilibrium(int nx, int ny, const double *__restrict__ u,
                              const double *__restrict__ rho, double *__restrict__ geq)
{
    const long long N = (long long)nx * ny;
    long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= N) return;
    const double ux = u[i], uy = u[N + i], r = rho[i];
    const double v = 1.5 * (ux * ux + uy * uy);
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        double t = 3.0 * (ux * (double)CX[q] + uy * (double)CY[q]);
        double e = 1.0 + t + 0.5 * (t * t) - v;
        geq[q * N + i] = e * (r * weight<double>(q));
    }
}

// nb.py:24-35: TRT collision into g_up, and g[0] = g_up[0].
__global__ void k_collide(int nx, int ny, double *__restrict__ g,
                          const double *__restrict__ geq, double *__restrict__ gup,
                          double om_p, double a1, double a2, double a3, double a4)
{
    const long long N = (long long)nx * ny;
    long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= N) return;
    double f[9], e[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        f[q] = g[q * N + i];
        e[q] = geq[q * N + i];
    }
    double up0 = (1.0 - om_p) * f[0] + om_p * e[0];
    gup[i] = up0;
    g[i] = up0;
#pragma unroll
    for (int q = 1; q < 9; ++q) {
        const int qb = NS[q];
        gup[q * N + i] = a1 * f[q] - a2 * f[qb] + a3 * e[q] + a4 * e[qb];
    }
}

// nb.py:37-45: push streaming, interior targets only; the other entries of g
// keep their values.
__global__ void k_stream(int nx, int ny, double *__restrict__ g,
                         const double *__restrict__ gup)
{
    const long long N = (long long)nx * ny;
    long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= N) return;
    const int x = (int)(i / ny), y = (int)(i - (long long)x * ny);
#pragma unroll
    for (int q = 1; q < 9; ++q) {
        const int sx = x - CX[q], sy = y - CY[q];
        if (sx < 0 || sx >= nx || sy < 0 || sy >= ny) continue;
        g[q * N + i] = gup[q * N + (long long)sx * ny + sy];
    }
}

__device__ inline bool wrap_index(long long k, int n, long long &out)
{
    if (k < -(long long)n || k >= n) return false;
    out = k < 0 ? k + n : k;
    return true;
}

// nb.py:49-73: sequential sums over the links (one thread, reference order).
__global__ void k_drag_sums(int nx, int ny, int nb, const int64_t *__restrict__ bnd,
                            const double *__restrict__ gup, const double *__restrict__ g,
                            double *__restrict__ out2)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const long long N = (long long)nx * ny;
    double fx = 0.0, fy = 0.0;
    for (int k = 0; k < nb; ++k) {
        long long i, j;
        if (!wrap_index(bnd[3 * k], nx, i) || !wrap_index(bnd[3 * k + 1], ny, j)) continue;
        const int q = (int)bnd[3 * k + 2];
        if (q < 0 || q > 8) continue;
        const int qb = NS[q];
        const long long c = i * ny + j;
        double g0 = gup[q * N + c] + g[qb * N + c];
        fx = fx + g0 * (double)CX[q];
        fy = fy + g0 * (double)CY[q];
    }
    out2[0] = fx;
    out2[1] = fy;
}

// nb.py:77-117 (compiled without fastmath in the reference).
__global__ void k_bounce_back(int nx, int ny, int ibb_on, int nb,
                              const int64_t *__restrict__ bnd, const double *__restrict__ pibb,
                              const double *__restrict__ gup, double *__restrict__ g)
{
    const long long N = (long long)nx * ny;
    int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= nb) return;
    long long i, j;
    if (!wrap_index(bnd[3 * k], nx, i) || !wrap_index(bnd[3 * k + 1], ny, j)) return;
    const int q = (int)bnd[3 * k + 2];
    if (q < 0 || q > 8) return;
    const int qb = NS[q];
    const long long c = i * ny + j;
    if (ibb_on) {
        long long im, jm, imm, jmm;
        if (!wrap_index(bnd[3 * k] + CX[qb], nx, im) ||
            !wrap_index(bnd[3 * k + 1] + CY[qb], ny, jm) ||
            !wrap_index(bnd[3 * k] + 2 * CX[qb], nx, imm) ||
            !wrap_index(bnd[3 * k + 1] + 2 * CY[qb], ny, jmm))
            return;
        const double p = pibb[k];
        const double pp = 2.0 * p;
        if (p < 0.5) {
            g[qb * N + c] = p * (pp + 1.0) * gup[q * N + c]
                          + (1.0 + pp) * (1.0 - pp) * gup[q * N + im * ny + jm]
                          - p * (1.0 - pp) * gup[q * N + imm * ny + jmm];
        } else {
            g[qb * N + c] = (1.0 / (p * (pp + 1.0))) * gup[q * N + c]
                          + ((pp - 1.0) / p) * gup[qb * N + c]
                          + ((1.0 - pp) / (1.0 + pp)) * gup[qb * N + im * ny + jm];
        }
    } else {
        g[qb * N + c] = gup[q * N + c];
    }
}

constexpr double CST1 = 2.0 / 3.0, CST2 = 1.0 / 6.0, CST3 = 1.0 / 2.0;

#define GQ(q, c) g[(long long)(q) * N + (c)]

// nb.py:121-143
__global__ void k_zh_left(int nx, int ny, double *u, const double *ul, double *rho, double *g)
{
    const long long N = (long long)nx * ny;
    int j = blockIdx.x * BLOCK + threadIdx.x;
    if (j >= ny) return;
    const long long c = j;  // i = 0
    const double ux = ul[j], uy = ul[ny + j];
    u[c] = ux;
    u[N + c] = uy;
    const double r = (GQ(0, c) + GQ(3, c) + GQ(4, c) + 2.0 * GQ(2, c) + 2.0 * GQ(6, c) +
                      2.0 * GQ(7, c)) / (1.0 - ux);
    rho[c] = r;
    GQ(1, c) = GQ(2, c) + CST1 * r * ux;
    GQ(5, c) = GQ(6, c) - CST3 * (GQ(3, c) - GQ(4, c)) + CST2 * r * ux + CST3 * r * uy;
    GQ(8, c) = GQ(7, c) + CST3 * (GQ(3, c) - GQ(4, c)) + CST2 * r * ux - CST3 * r * uy;
}

// nb.py:147-169
__global__ void k_zh_right_vel(int nx, int ny, double *u, const double *ur, double *rho, double *g)
{
    const long long N = (long long)nx * ny;
    int j = blockIdx.x * BLOCK + threadIdx.x;
    if (j >= ny) return;
    const long long c = (long long)(nx - 1) * ny + j;
    const double ux = ur[j], uy = ur[ny + j];
    u[c] = ux;
    u[N + c] = uy;
    const double r = (GQ(0, c) + GQ(3, c) + GQ(4, c) + 2.0 * GQ(1, c) + 2.0 * GQ(5, c) +
                      2.0 * GQ(8, c)) / (1.0 + ux);
    rho[c] = r;
    GQ(2, c) = GQ(1, c) - CST1 * r * ux;
    GQ(6, c) = GQ(5, c) + CST3 * (GQ(3, c) - GQ(4, c)) - CST2 * r * ux - CST3 * r * uy;
    GQ(7, c) = GQ(8, c) - CST3 * (GQ(3, c) - GQ(4, c)) - CST2 * r * ux + CST3 * r * uy;
}

// nb.py:173-195
__global__ void k_zh_right_pres(int nx, int ny, double *u, const double *rr, const double *ur,
                                double *rho, double *g)
{
    const long long N = (long long)nx * ny;
    int j = blockIdx.x * BLOCK + threadIdx.x;
    if (j >= ny) return;
    const long long c = (long long)(nx - 1) * ny + j;
    const double r = rr[j];
    rho[c] = r;
    const double uy = ur[ny + j];
    u[N + c] = uy;
    const double ux = (GQ(0, c) + GQ(3, c) + GQ(4, c) + 2.0 * GQ(1, c) + 2.0 * GQ(5, c) +
                       2.0 * GQ(8, c)) / r - 1.0;
    u[c] = ux;
    GQ(2, c) = GQ(1, c) - CST1 * r * ux;
    GQ(6, c) = GQ(5, c) + CST3 * (GQ(3, c) - GQ(4, c)) - CST2 * r * ux - CST3 * r * uy;
    GQ(7, c) = GQ(8, c) - CST3 * (GQ(3, c) - GQ(4, c)) - CST2 * r * ux + CST3 * r * uy;
}

// nb.py:199-221
__global__ void k_zh_top(int nx, int ny, double *u, const double *ut, double *rho, double *g)
{
    const long long N = (long long)nx * ny;
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nx) return;
    const long long c = (long long)i * ny + (ny - 1);
    const double ux = ut[i], uy = ut[nx + i];
    u[c] = ux;
    u[N + c] = uy;
    const double r = (GQ(0, c) + GQ(1, c) + GQ(2, c) + 2.0 * GQ(3, c) + 2.0 * GQ(5, c) +
                      2.0 * GQ(7, c)) / (1.0 + uy);
    rho[c] = r;
    GQ(4, c) = GQ(3, c) - CST1 * r * uy;
    GQ(8, c) = GQ(7, c) - CST3 * (GQ(1, c) - GQ(2, c)) + CST3 * r * ux - CST2 * r * uy;
    GQ(6, c) = GQ(5, c) + CST3 * (GQ(1, c) - GQ(2, c)) - CST3 * r * ux - CST2 * r * uy;
}

// nb.py:225-247
__global__ void k_zh_bottom(int nx, int ny, double *u, const double *ub, double *rho, double *g)
{
    const long long N = (long long)nx * ny;
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nx) return;
    const long long c = (long long)i * ny;
    const double ux = ub[i], uy = ub[nx + i];
    u[c] = ux;
    u[N + c] = uy;
    const double r = (GQ(0, c) + GQ(1, c) + GQ(2, c) + 2.0 * GQ(4, c) + 2.0 * GQ(6, c) +
                      2.0 * GQ(8, c)) / (1.0 - uy);
    rho[c] = r;
    GQ(3, c) = GQ(4, c) + CST1 * r * uy;
    GQ(5, c) = GQ(6, c) - CST3 * (GQ(1, c) - GQ(2, c)) + CST3 * r * ux + CST2 * r * uy;
    GQ(7, c) = GQ(8, c) + CST3 * (GQ(1, c) - GQ(2, c)) - CST3 * r * ux + CST2 * r * uy;
}

__device__ inline double corner_rest(const double *g, long long N, long long c, double r)
{
    double s = r;
#pragma unroll
    for (int q = 1; q < 9; ++q) s = s - GQ(q, c);
    return s;
}

// nb.py:251-344.  which: 0 bottom-left, 1 top-left, 2 top-right, 3 bottom-right.
__global__ void k_zh_corner(int nx, int ny, int which, double *u, double *rho, double *g)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const long long N = (long long)nx * ny;
    const int lx = nx - 1, ly = ny - 1;
    int i, j, ni;
    switch (which) {
    case 0: i = 0; j = 0; ni = 1; break;
    case 1: i = 0; j = ly; ni = 1; break;
    case 2: i = lx; j = ly; ni = lx - 1; break;
    default: i = lx; j = 0; ni = lx - 1; break;
    }
    const long long c = (long long)i * ny + j, cn = (long long)ni * ny + j;
    u[c] = u[cn];
    u[N + c] = u[N + cn];
    rho[c] = rho[cn];
    const double r = rho[c], ux = u[c], uy = u[N + c];
    switch (which) {
    case 0:
        GQ(1, c) = GQ(2, c) + (2.0 / 3.0) * r * ux;
        GQ(3, c) = GQ(4, c) + (2.0 / 3.0) * r * uy;
        GQ(5, c) = GQ(6, c) + (1.0 / 6.0) * r * ux + (1.0 / 6.0) * r * uy;
        GQ(7, c) = 0.0;
        GQ(8, c) = 0.0;
        break;
    case 1:
        GQ(1, c) = GQ(2, c) + (2.0 / 3.0) * r * ux;
        GQ(4, c) = GQ(3, c) - (2.0 / 3.0) * r * uy;
        GQ(8, c) = GQ(7, c) + (1.0 / 6.0) * r * ux - (1.0 / 6.0) * r * uy;
        GQ(5, c) = 0.0;
        GQ(6, c) = 0.0;
        break;
    case 2:
        GQ(2, c) = GQ(1, c) - (2.0 / 3.0) * r * ux;
        GQ(4, c) = GQ(3, c) - (2.0 / 3.0) * r * uy;
        GQ(6, c) = GQ(5, c) - (1.0 / 6.0) * r * ux - (1.0 / 6.0) * r * uy;
        GQ(7, c) = 0.0;
        GQ(8, c) = 0.0;
        break;
    default:
        GQ(2, c) = GQ(1, c) - (2.0 / 3.0) * r * ux;
        GQ(3, c) = GQ(4, c) + (2.0 / 3.0) * r * uy;
        GQ(7, c) = GQ(8, c) - (1.0 / 6.0) * r * ux + (1.0 / 6.0) * r * uy;
        GQ(5, c) = 0.0;
        GQ(6, c) = 0.0;
        break;
    }
    GQ(0, c) = corner_rest(g, N, c, r);
}

#undef GQ

int check_grid(int nx, int ny)
{
    if (nx < 3 || ny < 3) return set_error(LBM_ERR_ARG, "nx and ny must be >= 3");
    if ((long long)nx * ny * 9 >= (1LL << 40))
        return set_error(LBM_ERR_ARG, "lattice too large");
    return LBM_OK;
}

inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace
}  // namespace lbm

using namespace lbm;

extern "C" int lbm_macro(int nx, int ny, const double *g, double *rho, double *u, void *stream)
{
    if (int e = check_grid(nx, ny)) return e;
    if (!g || !rho || !u) return set_error(LBM_ERR_ARG, "lbm_macro: null pointer");
    k_macro<<<nblocks((long long)nx * ny), BLOCK, 0, S(stream)>>>(nx, ny, g, rho, u);
    LBM_LAUNCH_CHECK();
    return LBM_OK;
}

extern "C" int lbm_equilibrium(int nx, int ny, const double *u, const double *rho,
                               double *geq, void *stream)
{
    if (int e = check_grid(nx, ny)) return e;
    if (!geq || !rho || !u) return set_error(LBM_ERR_ARG, "lbm_equilibrium: null pointer");
    k_equilibrium<<<nblocks((long long)nx * ny), BLOCK, 0, S(stream)>>>(nx, ny, u, rho, geq);
    LBM_LAUNCH_CHECK();
    return LBM_OK;
}

extern "C" int lbm_col_str(int nx, int ny, double *g, const double *geq, double *gup,
                           double om_p, double om_m, void *stream)
{
    if (int e = check_grid(nx, ny)) return e;
    if (!g || !geq || !gup) return set_error(LBM_ERR_ARG, "lbm_col_str: null pointer");
    // the scalar factors of nb.py:32-35, formed once as the reference does
    const double a1 = 1.0 - 0.5 * (om_p + om_m);
    const double a2 = 0.5 * (om_p - om_m);
    const double a3 = 0.5 * (om_p + om_m);
    const double a4 = 0.5 * (om_p - om_m);
    const int nbk = nblocks((long long)nx * ny);
    k_collide<<<nbk, BLOCK, 0, S(stream)>>>(nx, ny, g, geq, gup, om_p, a1, a2, a3, a4);
    LBM_LAUNCH_CHECK();
    k_stream<<<nbk, BLOCK, 0, S(stream)>>>(nx, ny, g, gup);
    LBM_LAUNCH_CHECK();
    return LBM_OK;
}

extern "C" int lbm_drag_lift(int nx, int ny, int nb, const int64_t *bnd, const double *gup,
                             const double *g, double *out2, void *stream)
{
    if (int e = check_grid(nx, ny)) return e;
    if (nb < 0 || (nb > 0 && !bnd) || !gup || !g || !out2)
        return set_error(LBM_ERR_ARG, "lbm_drag_lift: bad argument");
    k_drag_sums<<<1, 64, 0, S(stream)>>>(nx, ny, nb, bnd, gup, g, out2);
    LBM_LAUNCH_CHECK();
    return LBM_OK;
}

extern "C" int lbm_bounce_back_obstacle(int nx, int ny, int ibb_on, int nb, const int64_t *bnd,
                                        const double *pibb, const double *gup, double *g,
                                        void *stream)
{
    if (int e = check_grid(nx, ny)) return e;
    if (nb < 0 || (nb > 0 && (!bnd || (ibb_on && !pibb))) || !gup || !g)
        return set_error(LBM_ERR_ARG, "lbm_bounce_back_obstacle: bad argument");
    if (nb == 0) return LBM_OK;
    k_bounce_back<<<nblocks(nb), BLOCK, 0, S(stream)>>>(nx, ny, ibb_on, nb, bnd, pibb, gup, g);
    LBM_LAUNCH_CHECK();
    return LBM_OK;
}

#define WALL_ENTRY(name, kern, len)                                                    \
    extern "C" int name(int nx, int ny, double *u, const double *us, double *rho,      \
                        double *g, void *stream)                                       \
    {                                                                                  \
        if (int e = check_grid(nx, ny)) return e;                                      \
        if (!u || !us || !rho || !g) return set_error(LBM_ERR_ARG, #name ": null");     \
        kern<<<nblocks(len), BLOCK, 0, S(stream)>>>(nx, ny, u, us, rho, g);            \
        LBM_LAUNCH_CHECK();                                                            \
        return LBM_OK;                                                                 \
    }

WALL_ENTRY(lbm_zou_he_left_wall_velocity, k_zh_left, ny)
WALL_ENTRY(lbm_zou_he_right_wall_velocity, k_zh_right_vel, ny)
WALL_ENTRY(lbm_zou_he_top_wall_velocity, k_zh_top, nx)
WALL_ENTRY(lbm_zou_he_bottom_wall_velocity, k_zh_bottom, nx)

extern "C" int lbm_zou_he_right_wall_pressure(int nx, int ny, double *u, const double *rr,
                                              const double *ur, double *rho, double *g,
                                              void *stream)
{
    if (int e = check_grid(nx, ny)) return e;
    if (!u || !rr || !ur || !rho || !g)
        return set_error(LBM_ERR_ARG, "lbm_zou_he_right_wall_pressure: null");
    k_zh_right_pres<<<nblocks(ny), BLOCK, 0, S(stream)>>>(nx, ny, u, rr, ur, rho, g);
    LBM_LAUNCH_CHECK();
    return LBM_OK;
}

#define CORNER_ENTRY(name, which)                                                      \
    extern "C" int name(int nx, int ny, double *u, double *rho, double *g, void *stream) \
    {                                                                                  \
        if (int e = check_grid(nx, ny)) return e;                                      \
        if (!u || !rho || !g) return set_error(LBM_ERR_ARG, #name ": null");            \
        k_zh_corner<<<1, 64, 0, S(stream)>>>(nx, ny, which, u, rho, g);                \
        LBM_LAUNCH_CHECK();                                                            \
        return LBM_OK;                                                                 \
    }

CORNER_ENTRY(lbm_zou_he_bottom_left_corner, 0)
CORNER_ENTRY(lbm_zou_he_top_left_corner, 1)
CORNER_ENTRY(lbm_zou_he_top_right_corner, 2)
CORNER_ENTRY(lbm_zou_he_bottom_right_corner, 3)
