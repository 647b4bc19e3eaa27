/* Sanitizer driver for the C oracle (test infrastructure): built together
 * with oracle/d2q9_oracle.c under -fsanitize=address,undefined by
 * tests/test_oracle_sanitize.py and run as a child process.  It drives the
 * reference's call order (lbm/src/core/run.py:24-54, turek.py:112-125) on a
 * small channel with a square obstacle whose links carry interpolated
 * bounce-back distances on both sides of 1/2, links next to the lattice edge
 * (the wrapped indices of the step app) and the four Zou-He walls and
 * corners, then checks the error returns of out-of-range links.  Exit status
 * 0 with "ok" on stdout; any invalid access aborts under the sanitizers. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

void oracle_macro(int nx, int ny, const double *g, double *rho, double *u);
void oracle_equilibrium(int nx, int ny, const double *u, const double *rho, double *g_eq);
void oracle_col_str(int nx, int ny, double *g, const double *g_eq, double *g_up,
                    double om_p, double om_m);
int oracle_drag_sums(int nx, int ny, int nb, const int64_t *boundary, const double *g_up,
                     const double *g, double *out2);
int oracle_bounce_back_obstacle(int nx, int ny, int ibb_on, int nb, const int64_t *boundary,
                                const double *obs_ibb, const double *g_up, double *g);
void oracle_zou_he_left_wall_velocity(int nx, int ny, double *u, const double *u_left,
                                      double *rho, double *g);
void oracle_zou_he_right_wall_pressure(int nx, int ny, double *u, const double *rho_right,
                                       const double *u_right, double *rho, double *g);
void oracle_zou_he_top_wall_velocity(int nx, int ny, double *u, const double *u_top,
                                     double *rho, double *g);
void oracle_zou_he_bottom_wall_velocity(int nx, int ny, double *u, const double *u_bot,
                                        double *rho, double *g);
void oracle_zou_he_bottom_left_corner(int nx, int ny, double *u, double *rho, double *g);
void oracle_zou_he_top_left_corner(int nx, int ny, double *u, double *rho, double *g);
void oracle_zou_he_top_right_corner(int nx, int ny, double *u, double *rho, double *g);
void oracle_zou_he_bottom_right_corner(int nx, int ny, double *u, double *rho, double *g);

static const int CX[9] = {0, 1, -1, 0, 0, 1, -1, -1, 1};
static const int CY[9] = {0, 0, 0, 1, -1, 1, -1, 1, -1};

static int solid(int i, int j)
{
    /* a 6x6 square, and a 2x3 block touching the bottom wall */
    return (i >= 12 && i < 18 && j >= 7 && j < 13) || (i >= 30 && i < 32 && j < 3);
}

int main(void)
{
    const int nx = 48, ny = 20, n = nx * ny;
    double *g = calloc((size_t)9 * n, sizeof(double));
    double *g_eq = calloc((size_t)9 * n, sizeof(double));
    double *g_up = calloc((size_t)9 * n, sizeof(double));
    double *rho = calloc((size_t)n, sizeof(double));
    double *u = calloc((size_t)2 * n, sizeof(double));
    double *u_left = calloc((size_t)2 * ny, sizeof(double));
    double *u_right = calloc((size_t)2 * ny, sizeof(double));
    double *u_top = calloc((size_t)2 * nx, sizeof(double));
    double *u_bot = calloc((size_t)2 * nx, sizeof(double));
    double *rho_right = calloc((size_t)ny, sizeof(double));
    int64_t *links = malloc(sizeof(int64_t) * 3 * 9 * n);
    double *dist = malloc(sizeof(double) * 9 * n);
    if (!g || !g_eq || !g_up || !rho || !u || !u_left || !u_right || !u_top || !u_bot ||
        !rho_right || !links || !dist)
        return 2;
    /* fluid nodes next to the solid: link (i, j, q) with q pointing into it;
     * distances alternate below and above 1/2 (both IBB branches); links of
     * the bottom block reach j - 2 < 0 and wrap, as the reference's indices do */
    int nb = 0;
    for (int i = 1; i < nx - 1; ++i)
        for (int j = 0; j < ny; ++j) {
            if (solid(i, j)) continue;
            for (int q = 1; q < 9; ++q) {
                const int si = i + CX[q], sj = j + CY[q];
                if (si < 0 || si >= nx || sj < 0 || sj >= ny || !solid(si, sj)) continue;
                links[3 * nb + 0] = i;
                links[3 * nb + 1] = j;
                links[3 * nb + 2] = q;
                dist[nb] = (nb % 2) ? 0.3 + 0.01 * (nb % 7) : 0.55 + 0.03 * (nb % 11);
                ++nb;
            }
        }
    for (int j = 0; j < ny; ++j) {
        u_left[j] = 0.04 * 4.0 * j * (ny - 1 - j) / ((ny - 1.0) * (ny - 1.0));
        rho_right[j] = 1.0;
    }
    for (int k = 0; k < n; ++k) rho[k] = 1.0;
    oracle_equilibrium(nx, ny, u, rho, g);
    double drag[2] = {0.0, 0.0};
    for (int it = 0; it < 200; ++it) {
        oracle_macro(nx, ny, g, rho, u);
        oracle_equilibrium(nx, ny, u, rho, g_eq);
        oracle_col_str(nx, ny, g, g_eq, g_up, 1.0 / 0.56, 1.0 / (0.25 / 0.06 + 0.5));
        if (oracle_bounce_back_obstacle(nx, ny, 1, nb, links, dist, g_up, g)) return 3;
        oracle_zou_he_left_wall_velocity(nx, ny, u, u_left, rho, g);
        oracle_zou_he_right_wall_pressure(nx, ny, u, rho_right, u_right, rho, g);
        oracle_zou_he_top_wall_velocity(nx, ny, u, u_top, rho, g);
        oracle_zou_he_bottom_wall_velocity(nx, ny, u, u_bot, rho, g);
        oracle_zou_he_bottom_left_corner(nx, ny, u, rho, g);
        oracle_zou_he_top_left_corner(nx, ny, u, rho, g);
        oracle_zou_he_top_right_corner(nx, ny, u, rho, g);
        oracle_zou_he_bottom_right_corner(nx, ny, u, rho, g);
        if (oracle_drag_sums(nx, ny, nb, links, g_up, g, drag)) return 4;
    }
    for (int k = 0; k < 9 * n; ++k)
        if (!isfinite(g[k])) return 5;
    /* plain bounce-back on the same links */
    if (oracle_bounce_back_obstacle(nx, ny, 0, nb, links, dist, g_up, g)) return 6;
    /* out-of-range links are refused with an error code, not read */
    int64_t bad[3] = {nx + 3, 2, 1};
    if (oracle_bounce_back_obstacle(nx, ny, 1, 1, bad, dist, g_up, g) >= 0) return 7;
    if (oracle_drag_sums(nx, ny, 1, bad, g_up, g, drag) >= 0) return 8;
    int64_t edge[3] = {1, 0, 3};   /* IBB stencil reaching j = -1, -2 below the lower edge wraps */
    if (oracle_bounce_back_obstacle(nx, ny, 1, 1, edge, dist, g_up, g)) return 9;
    int64_t far[3] = {0, 2, 1};    /* IBB stencil reaching i = -1, -2 wraps */
    if (oracle_bounce_back_obstacle(nx, ny, 1, 1, far, dist, g_up, g)) return 10;
    int64_t beyond[3] = {-nx - 1, 2, 2};   /* below -nx: refused */
    if (oracle_bounce_back_obstacle(nx, ny, 1, 1, beyond, dist, g_up, g) >= 0) return 11;
    printf("ok %d links, drag %.6e %.6e\n", nb, drag[0], drag[1]);
    free(g); free(g_eq); free(g_up); free(rho); free(u); free(u_left); free(u_right);
    free(u_top); free(u_bot); free(rho_right); free(links); free(dist);
    return 0;
}
