/*
 * lbm_d2q9.h -- C ABI of the MI355X (gfx950) D2Q9 lattice-Boltzmann core.
 *
 * Library: lbm_amd/liblbm_d2q9.so (hand-written HIP, built by
 * lbm_amd/csrc/Makefile).  Every entry point returns an int status
 * (LBM_OK == 0); lbm_last_error() gives the message of the last failure on
 * the calling thread.  Pointers named *_dev are device pointers (HBM);
 * pointers named *_host are host pointers.  `stream` is a hipStream_t passed
 * as void* (NULL = the default stream).  Nothing in this header depends on
 * torch: a C, ctypes or cgo caller binds it directly.
 *
 * Data layout (same as the reference, lbm/src/core/lattice.py:155-174):
 *   populations g, g_eq, g_up : double[9][nx][ny], y fastest
 *   density rho               : double[nx][ny]
 *   velocity u                : double[2][nx][ny]
 *   wall inputs u_left/u_right: double[2][ny]; u_top/u_bot: double[2][nx];
 *   rho_right                 : double[ny]
 *   obstacle links "boundary" : int64[nb][3] = (i, j, q), q pointing from the
 *                               fluid node towards the solid (lattice.py:328-344)
 * The velocity set c, the weights w and the bounce-back table ns are the
 * fixed D2Q9 tables of lattice.py:135-152; they are not arguments.
 */
#ifndef LBM_D2Q9_H
#define LBM_D2Q9_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_OK              0
#define LBM_ERR_ARG        -1   /* bad size, null pointer, index out of range */
#define LBM_ERR_HIP        -2   /* a HIP runtime call failed */
#define LBM_ERR_UNSUPPORTED -3  /* a boundary-condition plan the fused path cannot express */

#define LBM_ABI_VERSION     1

/* Population storage of the fused path. */
#define LBM_STORE_F64   0   /* double storage, double arithmetic */
#define LBM_STORE_F32   1   /* float storage of f_q - w_q, float arithmetic */
#define LBM_STORE_F16   2   /* IEEE half storage of f_q - w_q, float arithmetic */
#define LBM_STORE_BF16  3   /* bfloat16 storage of f_q - w_q, float arithmetic */
#define LBM_STORE_F16E  4   /* IEEE half storage of f_q - w_q - E_q(v), E_q the
                               equilibrium deviation at the node's velocity v,
                               which a tenth plane keeps as a half pair; float
                               arithmetic (fused and batched paths) */
#define LBM_STORE_F16E8 6   /* as LBM_STORE_F16E with the velocity pair as 8-bit
                               fixed point (steps of 2^-9, |v| <= 127/512) in a
                               16-bit tenth plane: 40 bytes per update */
#define LBM_STORE_Q16   5   /* 16-bit fixed point of f_q - w_q on a per-direction
                               affine map covering a caller-given range (the
                               q16_range fields); float arithmetic (fused and
                               batched paths) */

/* Wall kinds of the fused boundary-condition plan. */
#define LBM_BC_NONE      0
#define LBM_BC_VELOCITY  1   /* nb_zou_he_*_wall_velocity */
#define LBM_BC_PRESSURE  2   /* nb_zou_he_right_wall_pressure (right wall only) */

/* Sides, in the order of the inlet table rows. */
#define LBM_SIDE_BOTTOM     0   /* u_bot    [2][nx] */
#define LBM_SIDE_LEFT       1   /* u_left   [2][ny] */
#define LBM_SIDE_TOP        2   /* u_top    [2][nx] */
#define LBM_SIDE_RIGHT      3   /* u_right  [2][ny] */
#define LBM_SIDE_RHO_RIGHT  4   /* rho_right   [ny] */

int         lbm_abi_version(void);
const char *lbm_last_error(void);

/* ------------------------------------------------------------------------
 * Operation-by-operation kernels.  Each replaces one reference function with
 * the same arguments and in-place semantics (double precision, strict
 * evaluation order, bit-identical to a strict-IEEE evaluation of the
 * reference expression).  lx, ly of the reference are nx-1, ny-1.
 * ---------------------------------------------------------------------- */

/* lattice.macro, lbm/src/core/lattice.py:178-189 */
int lbm_macro(int nx, int ny, const double *g_dev, double *rho_dev, double *u_dev,
              void *stream);

/* nb_equilibrium(u, c, w, rho, g_eq), lbm/src/core/nb.py:7-17 */
int lbm_equilibrium(int nx, int ny, const double *u_dev, const double *rho_dev,
                    double *g_eq_dev, void *stream);

/* nb_col_str(g, g_eq, g_up, om_p, om_m, c, ns, nx, ny, lx, ly), nb.py:21-45 */
int lbm_col_str(int nx, int ny, double *g_dev, const double *g_eq_dev, double *g_up_dev,
                double om_p, double om_m, void *stream);

/* nb_drag_lift(boundary, ns, c, g_up, g, R_ref, U_ref, L_ref), nb.py:49-73.
 * Writes the momentum-exchange sums (fx, fy), accumulated over the links in
 * order, to out2_dev[0..1]; Cx = -2 fx / (R_ref L_ref U_ref^2) and Cy likewise
 * (nb.py:70-71) are formed by the caller. */
int lbm_drag_lift(int nx, int ny, int nb, const int64_t *boundary_dev,
                  const double *g_up_dev, const double *g_dev, double *out2_dev,
                  void *stream);

/* nb_bounce_back_obstacle(IBB, boundary, ns, c, ibb, g_up, g, u, lattice),
 * nb.py:77-117 (the unused u and lattice arguments are dropped). */
int lbm_bounce_back_obstacle(int nx, int ny, int ibb_on, int nb,
                             const int64_t *boundary_dev, const double *ibb_dev,
                             const double *g_up_dev, double *g_dev, void *stream);

/* nb_zou_he_left_wall_velocity(lx, ly, u, u_left, rho, g), nb.py:121-143 */
int lbm_zou_he_left_wall_velocity(int nx, int ny, double *u_dev, const double *u_left_dev,
                                  double *rho_dev, double *g_dev, void *stream);
/* nb_zou_he_right_wall_velocity, nb.py:147-169 */
int lbm_zou_he_right_wall_velocity(int nx, int ny, double *u_dev, const double *u_right_dev,
                                   double *rho_dev, double *g_dev, void *stream);
/* nb_zou_he_right_wall_pressure(lx, ly, u, rho_right, u_right, rho, g), nb.py:173-195 */
int lbm_zou_he_right_wall_pressure(int nx, int ny, double *u_dev, const double *rho_right_dev,
                                   const double *u_right_dev, double *rho_dev, double *g_dev,
                                   void *stream);
/* nb_zou_he_top_wall_velocity, nb.py:199-221 */
int lbm_zou_he_top_wall_velocity(int nx, int ny, double *u_dev, const double *u_top_dev,
                                 double *rho_dev, double *g_dev, void *stream);
/* nb_zou_he_bottom_wall_velocity, nb.py:225-247 */
int lbm_zou_he_bottom_wall_velocity(int nx, int ny, double *u_dev, const double *u_bot_dev,
                                    double *rho_dev, double *g_dev, void *stream);
/* nb_zou_he_bottom_left_corner_velocity, nb.py:251-271 */
int lbm_zou_he_bottom_left_corner(int nx, int ny, double *u_dev, double *rho_dev,
                                  double *g_dev, void *stream);
/* nb_zou_he_top_left_corner_velocity, nb.py:275-296 */
int lbm_zou_he_top_left_corner(int nx, int ny, double *u_dev, double *rho_dev,
                               double *g_dev, void *stream);
/* nb_zou_he_top_right_corner_velocity, nb.py:300-320 */
int lbm_zou_he_top_right_corner(int nx, int ny, double *u_dev, double *rho_dev,
                                double *g_dev, void *stream);
/* nb_zou_he_bottom_right_corner_velocity, nb.py:324-344 */
int lbm_zou_he_bottom_right_corner(int nx, int ny, double *u_dev, double *rho_dev,
                                   double *g_dev, void *stream);

/* ------------------------------------------------------------------------
 * Fused time step.  One launch of lbm_step performs, for every node, what
 * the reference does across one pass of its loop (run.py:24-54) between two
 * macro() calls: pull streaming of the previous post-collision populations,
 * the recorded boundary-condition plan (obstacle BB/IBB links, Zou-He walls,
 * Zou-He corners, in the reference's set_bc order), the drag/lift
 * momentum-exchange terms of the step just finished, the moments, the
 * equilibrium and the TRT collision.  The state between launches is the
 * post-collision field g_up, held by the handle in two ping-pong buffers in
 * the chosen storage format.
 *
 * Step numbering: iteration `it` of the reference loop is executed by
 * lbm_init (it == 0) or lbm_step (it >= 1) for its macro/equilibrium/
 * collision half, and by the next lbm_step (or lbm_epilogue) for its
 * stream/boundary/drag half.  Inlet values of iteration it and the drag of
 * iteration it live in row (it mod ring) of the inlet table and drag ring.
 * ---------------------------------------------------------------------- */
typedef struct lbm_fused lbm_fused;

typedef struct {
    int32_t nx, ny;
    int32_t storage;            /* LBM_STORE_* */
    double  om_p, om_m;         /* TRT rates 1/tau_p, 1/tau_m (lattice.py:127-131) */
    int32_t bc[4];              /* LBM_BC_* for bottom, left, top, right */
    int32_t corners;            /* 1: the four Zou-He corners follow the walls */
    int32_t obstacles_first;    /* 1: obstacle links are applied before the walls */
    int32_t n_links;            /* all links of all obstacles, in plan order */
    const int64_t *links_host;  /* [n_links][3] (i, j, q); negative i, j wrap */
    const double  *ibb_host;    /* [n_links] IBB distance p (ignored for plain BB) */
    const uint8_t *link_ibb_host; /* [n_links] 1 = interpolated (IBB), 0 = plain BB */
    int32_t n_obs;              /* obstacles (drag/lift sums are kept per obstacle) */
    const int32_t *obs_offsets_host; /* [n_obs+1] link range of each obstacle */
    int32_t ring;               /* rows of the inlet table and of the drag ring */
    int32_t cells_per_thread;   /* 1, 2, 4 or 8 (0 = library default) */
    int32_t vector_width;       /* 16-bit forms: 4 = four consecutive nodes per
                                   work-item with 8-byte accesses (taken when
                                   ny % 4 == 0 and nx*ny % (1024*cpt) == 0),
                                   0 or 1 = one node per work-item */
    const double *q16_range_host; /* LBM_STORE_Q16: [9][2] (lo, hi) of h_q = f_q - w_q
                                   per direction, -1 < lo < hi < 1; values outside
                                   are clamped.  Ignored by the other forms. */
} lbm_fused_desc;

int  lbm_fused_create(const lbm_fused_desc *desc, lbm_fused **out);
void lbm_fused_destroy(lbm_fused *h);

/* Inlet table.  Row r holds, for each of the five sides, an offset (in
 * doubles) into a pool of side arrays: u_bot [2][nx], u_left [2][ny], u_top
 * [2][nx], u_right [2][ny], rho_right [ny].  lbm_fused_set_inlets replaces
 * the pool (pool_len doubles) and writes nrows rows starting at row0 (mod
 * ring).  Every offset must leave its whole side array inside the pool
 * (offset + length <= pool_len), unused sides included; otherwise nothing is
 * written and LBM_ERR_ARG is returned. */
int lbm_fused_set_inlets(lbm_fused *h, int row0, int nrows, const int32_t *offsets_host,
                         int pool_len, const double *pool_host, void *stream);

/* Separable wall inputs (no counterpart in the reference, whose apps fill
 * the wall arrays in Python every iteration; turek.py:94-109,
 * cavity.py:65-76 are of this form): after lbm_fused_set_profiles, the
 * value of side s (LBM_SIDE_*), element i, at iteration it is
 * amp[it mod ring][s] * profile_s[i], and the inlet table of
 * lbm_fused_set_inlets is no longer read.  profile_host[s] is u_bot/u_top
 * [2][nx], u_left/u_right [2][ny] or rho_right [ny] (NULL = zeros);
 * lbm_fused_set_amplitudes writes rows row0 .. row0+nrows-1 (mod ring) of
 * amp_host[nrows][5].  Not for the cases of a batch (their amplitudes are
 * the batch's). */
int lbm_fused_set_profiles(lbm_fused *h, const double *const *profile_host, void *stream);
int lbm_fused_set_amplitudes(lbm_fused *h, int64_t row0, int nrows, const double *amp_host,
                             void *stream);

/* Iteration 0: load g (double[9][nx][ny] on the device; NULL = the rest
 * state g = w), take moments, collide, store the post-collision field.  When rho_dev/u_dev are not
 * NULL the moments of g (what lattice.macro() leaves) are stored there. */
int lbm_init(lbm_fused *h, const double *g_dev, double *rho_dev, double *u_dev, void *stream);

/* Iteration it >= 1 (it is the absolute iteration number). */
int lbm_step(lbm_fused *h, int64_t it, double *rho_dev, double *u_dev, void *stream);

/* Iterations it0 .. it0+n-1 (it0 >= 1) without moment output, replayed from
 * a captured HIP graph. */
int lbm_steps(lbm_fused *h, int64_t it0, int n, void *stream);

/* Captures, instantiates and uploads the graph lbm_steps(h, it0, n) replays
 * (one per length n and parity of it0) without running it, so that a later
 * lbm_steps call with that key launches only.  No counterpart in the
 * reference (it has no graphs); for callers that time or capture steps. */
int lbm_steps_prepare(lbm_fused *h, int64_t it0, int n, void *stream);

/* Stream/boundary/drag half of iteration it (the last one executed):
 * writes the post-boundary populations g (double[9][nx][ny]), the
 * post-collision field g_up of iteration it (double), and overwrites rho/u on
 * the domain perimeter with the values the Zou-He routines leave there. */
int lbm_epilogue(lbm_fused *h, int64_t it, double *g_dev, double *g_up_dev,
                 double *rho_dev, double *u_dev, void *stream);

/* Drag/lift terms of iteration it only (link nodes only, no state change). */
int lbm_drag_only(lbm_fused *h, int64_t it, void *stream);

/* Momentum-exchange sums of iterations it0 .. it0+n-1: out_host[n][n_obs][2]
 * = (fx, fy), each summed over the obstacle's links in plan order.
 * Synchronises the stream. */
int lbm_drag_read(lbm_fused *h, int64_t it0, int n, double *out_host, void *stream);

/* Allocation-free variant of lbm_drag_read.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as lbm_drag_read, but the per-obstacle sums are
 * staged in a fixed device array of the library (LBM_NOALLOC_DRAG_MAX
 * doubles, 8 MiB, one per GPU, held until the process ends) instead of a
 * buffer lbm_drag_read allocates and grows: n * n_obs * 2 must not exceed
 * LBM_NOALLOC_DRAG_MAX (else LBM_ERR_ARG and nothing is read).  Every call on
 * one GPU uses that GPU's array, so calls that overlap on one GPU (other
 * streams or host threads) race.  Synchronises the stream. */
#define LBM_NOALLOC_DRAG_MAX 1048576
int lbm_drag_read_noalloc(lbm_fused *h, int64_t it0, int n, double *out_host, void *stream);

/* Checkpoint of the post-collision state of iteration it (for rollback).
 * The handle keeps the two most recent checkpoints: a save replaces the
 * older one (or the one of the same iteration); restore takes either. */
int lbm_fused_save(lbm_fused *h, int64_t it, void *stream);
int lbm_fused_restore(lbm_fused *h, int64_t it, void *stream);

/* FP64 density-drift check: total mass sum_x sum_q g_up_q(x) of the current
 * post-collision state of iteration it, reduced in double on the device.
 * Synchronises the stream. */
int lbm_mass(lbm_fused *h, int64_t it, double *out_host, void *stream);

/* Bytes of device memory held by the handle. */
int64_t lbm_fused_device_bytes(const lbm_fused *h);

/* ------------------------------------------------------------------------
 * Observable histories (host code, no GPU; no counterpart in the reference's
 * compiled code): buff.mv_avg (lbm/src/utils/buff.py:51-77), the triple
 * moving average over the last fifth of a history, for k consecutive
 * iterations at once, with the values np.sum gives (blocked pairwise
 * summation).  The arrays are buff's: hist = the history, avg1..avg3 = the
 * three averages, each starting with two zeros.  Iterations it0+1 .. it0+k:
 * hist must hold entries 0 .. it0+k+1; entries it0+2 .. it0+k+1 of avg1..avg3
 * are written, and obs_out[t] receives the value mv_avg returns at iteration
 * it0+1+t.  Large chunks are split over up to 16 host threads.
 * ---------------------------------------------------------------------- */
int lbm_mv_avg_chunk(const double *hist, double *avg1, double *avg2, double *avg3,
                     int64_t it0, int k, double *obs_out);

/* ------------------------------------------------------------------------
 * Batched ensembles (no counterpart in the reference, which runs one case per
 * process): many lattices of different sizes, rates, walls and obstacles
 * advanced together, one launch per iteration for the whole batch (a grid-
 * stride sweep over every case's tiles), replayed from a HIP graph.  Each
 * case is a full lbm_fused handle (lbm_batch_case) for per-case reads.
 * Wall inputs are separable: the value of side s, element i, at iteration
 * it is amp[it][case][s] * profile_s[i] -- the reference apps' inlets are of
 * this form ((ret*u_lbm) * profile_j, cavity.py:65-76, turek.py:94-109).
 * ---------------------------------------------------------------------- */
typedef struct lbm_batch lbm_batch;

typedef struct {
    int32_t nx, ny;
    double  om_p, om_m;
    int32_t bc[4];                  /* as lbm_fused_desc.bc; corners always on */
    const double *profile_host[5];  /* per side (LBM_SIDE_*): [2][nx|ny], rho_right [ny]; NULL = 0 */
    int32_t n_links;                /* optional obstacle links, as lbm_fused_desc */
    const int64_t *links_host;
    const double  *ibb_host;
    const uint8_t *link_ibb_host;
    int32_t n_obs;
    const int32_t *obs_offsets_host;
} lbm_case_desc;

typedef struct {
    int32_t n_cases;
    const lbm_case_desc *cases;
    int32_t storage;                /* LBM_STORE_*, common to all cases */
    int32_t cells_per_thread;       /* tile shape, common to all cases (0 = default) */
    int32_t vector_width;           /* as lbm_fused_desc */
    int32_t ring;                   /* rows of the amplitude table and drag rings */
    int32_t grid;                   /* work-groups of the batched launch (0 = one per work item) */
    const double *q16_range_host;   /* LBM_STORE_Q16: as lbm_fused_desc, common to all cases */
} lbm_batch_desc;

int  lbm_batch_create(const lbm_batch_desc *desc, lbm_batch **out);
void lbm_batch_destroy(lbm_batch *b);
/* case c's handle (owned by the batch): lbm_step/lbm_epilogue/lbm_drag_read/... on one case */
int  lbm_batch_case(lbm_batch *b, int c, lbm_fused **out);
/* out4 = {work items, grid, cells per thread, vector width} */
int  lbm_batch_info(const lbm_batch *b, int32_t *out4);
/* amplitudes of iterations row0 .. row0+nrows-1: amp_host[nrows][n_cases][5] */
int  lbm_batch_set_amplitudes(lbm_batch *b, int64_t row0, int nrows, const double *amp_host,
                              void *stream);
/* iteration 0 of every case from the rest state g = w */
int  lbm_batch_init(lbm_batch *b, void *stream);
/* iterations it0 .. it0+n-1 (it0 >= 1) of every case, graph-replayed */
int  lbm_batch_steps(lbm_batch *b, int64_t it0, int n, void *stream);
/* the graph of lbm_batch_steps(b, it0, n), captured and uploaded, not run */
int  lbm_batch_steps_prepare(lbm_batch *b, int64_t it0, int n, void *stream);

/* ------------------------------------------------------------------------
 * Drag sensitivity by a discrete adjoint (no counterpart in the reference).
 * A channel with the channel apps' boundary plan (Zou-He walls at rest,
 * velocity inlet u = a(it) * profile on the left with a(it) = ramp(it) * u_max,
 * pressure outlet rho_out, the four corners; turek.py:94-125) and an obstacle
 * given by a per-node solid fraction s in [0,1] (partial bounce-back:
 * post-collision H = (1-s) c + s h_opposite).  The loss is the mean over the
 * last `window` iterations of the momentum-exchange drag
 * D = sum_x s sum_q c_qx (c_q - h_opposite).  lbm_pbb_forward keeps the state
 * every `checkpoint_every` iterations; lbm_pbb_backward recomputes each
 * segment from its checkpoint and runs the hand-written adjoint step
 * backwards, giving dL/ds (every node) and dL/du_max.  Storage: LBM_STORE_F64
 * or LBM_STORE_F32 (states and arithmetic); drag sums and gradient
 * accumulators are double.
 * ---------------------------------------------------------------------- */
typedef struct lbm_pbb lbm_pbb;

typedef struct {
    int32_t nx, ny;
    int32_t storage;                 /* LBM_STORE_F64 or LBM_STORE_F32 */
    double  om_p, om_m;              /* TRT rates */
    double  rho_out;                 /* outlet density */
    const double *inlet_profile_host;/* [2][ny] inlet velocity of unit amplitude */
    const double *ramp_host;         /* [iterations+1] ramp(it) */
    int32_t iterations;              /* T: iterations 1..T after the start from rest */
    int32_t window;                  /* the loss averages D over iterations T-window+1..T */
    int32_t checkpoint_every;
} lbm_pbb_desc;

int  lbm_pbb_create(const lbm_pbb_desc *desc, lbm_pbb **out);
void lbm_pbb_destroy(lbm_pbb *h);
/* s_dev: [nx][ny] solid fraction in the storage type; writes the loss */
int  lbm_pbb_forward(lbm_pbb *h, const void *s_dev, double u_max, double *loss_host, void *stream);
/* after lbm_pbb_forward: s_grad_dev [nx][ny] double, *umax_grad_host */
int  lbm_pbb_backward(lbm_pbb *h, double *s_grad_dev, double *umax_grad_host, void *stream);
/* drag D of iterations 0..T of the last forward run: out_host[T+1] */
int  lbm_pbb_drag_history(const lbm_pbb *h, double *out_host);
/* post-collision deviations H_T = f - w of the last forward run ([9][nx][ny], storage type) */
int  lbm_pbb_final_state(lbm_pbb *h, void *dst_dev, void *stream);

/* ------------------------------------------------------------------------
 * Measurement utility (no counterpart in the reference): device-to-device
 * copy of `bytes` (a multiple of 16) with `blocks` work-groups of 256
 * work-items and 16-byte accesses -- the HBM ceiling the time step is
 * compared with, and the kernel of known traffic that calibrates the
 * FETCH_SIZE/WRITE_SIZE counters.
 * ---------------------------------------------------------------------- */
int lbm_copy_probe(void *dst_dev, const void *src_dev, int64_t bytes, int blocks, void *stream);
/* the same with non-temporal 16-byte loads and stores */
int lbm_copy_probe_nt(void *dst_dev, const void *src_dev, int64_t bytes, int blocks,
                      void *stream);

#ifdef __cplusplus
}
#endif
#endif /* LBM_D2Q9_H */
