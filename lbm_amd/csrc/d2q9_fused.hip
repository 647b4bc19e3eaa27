// d2q9_fused.hip -- fused pull-scheme D2Q9 time step for gfx950 (MI355X).
//
// One launch per time step.  Each work-item owns a node and
//   1. pulls its nine populations from the previous post-collision field
//      (streaming, lbm/src/core/nb.py:37-45, read as a gather),
//   2. on obstacle-link nodes, replaces the populations the reference's
//      nb_bounce_back_obstacle writes (nb.py:77-117),
//   3. on perimeter nodes, applies the Zou-He wall or corner routine
//      (nb.py:121-344) in registers,
//   4. on link nodes, stores the momentum-exchange term of nb_drag_lift
//      (nb.py:64) into a ring row, summed later in link order,
//   5. takes moments (lattice.py:178-189), forms the equilibrium
//      (nb.py:7-17) and collides (nb.py:24-35), then stores the nine
//      post-collision populations.
// Interior nodes of tiles without links take a branch-free path: nine
// coalesced loads at constant offsets, arithmetic, nine coalesced stores.
// HBM traffic per node and step is 9 loads + 9 stores of the storage type.
//
// Storage forms: double; float; IEEE half or bfloat16 holding f_q - w_q
// (float arithmetic in both).  Drag terms are accumulated in double.
#include "d2q9_common.h"
#include "d2q9_zouhe.h"
#include "../../include/lbm_d2q9.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <vector>

namespace lbm {

thread_local std::string g_last_error;

int set_error(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

// ------------------------------------------------------------------ links --
// One obstacle link (fluid node `cell`, direction q towards the solid), with
// the reads of its bounce-back value resolved to (plane, node) pairs and the
// interpolation factors of nb.py:97-104 formed on the host in double.
struct LinkRec {
    int32_t cell;          // fluid node, x*ny + y
    int32_t cellB, cellC;  // nodes of the 2nd and 3rd reads
    int8_t q, qb, planeB, planeC;
    int32_t mode;          // 0 plain bounce-back, 1 interpolated
    int32_t dslot;         // position of the link in plan order (drag row)
    double c0, c1, c2;     // value = (c0*A + c1*B) + c2*C
};

struct Params {
    const void *src;       // post-collision field of iteration it-1 (storage form)
    void *dst;             // post-collision field of iteration it
    const double *g0;      // INIT: populations of iteration 0, double[9][N]
    double *g_out;         // EPILOGUE: post-boundary populations, double[9][N]
    double *gup_out;       // EPILOGUE: src converted to double[9][N]
    double *rho_out, *u_out;
    int nx, ny, N;
    uint32_t magic;        // floor(2^32 / ny) + 1, for node -> (x, y)
    double om_p, om_m, b0, a1, a2, a3, a4;
    float omf_p, omf_m;    // om_p and om_m rounded to float (the float forms' rates)
    int bc_bottom, bc_left, bc_top, bc_right, corners;
    const double *pool;    // inlet side arrays
    const int *inlet_off;  // [ring][5] offsets into pool
    const double *amp;     // separable inlets: amp[row * amp_stride + side] (or null)
    int amp_stride;
    const double *prof[5]; // separable inlets: per-side profiles
    int ring;
    const int *base;       // device int: row of iteration (it0-1) in graph replays
    int k;                 // row = (*base + k) mod ring  -- row of iteration it-1
    const int *tile_slot;  // [n_tiles] -1 or slot of the tile's link offsets
    const int *cell_off;   // [n_slots][tile_cells+1]
    const uint32_t *cell_mask; // [n_slots][tile_cells/32] bit set = node has links
    const int *wave_rng;   // float forms: [n_tiles][rng_stride] link-record bounds of
    int rng_stride;        // the tile's 64-node runs (every tile; equal bounds: no links)
    const uint32_t *run_bits; // float forms: bit r set = 64-node run r (node >> 6) has links
    int n_main;            // work-groups of interior tiles (the rest do the perimeter)
    int n_perim;           // perimeter nodes
    const LinkRec *links;  // sorted by node
    double *drag_ring;     // [ring][n_links]
    int n_links;
    const int *link_cells; // DRAG: distinct link nodes
    int n_link_cells;
    int tile_shift;        // log2(tile_cells)
    int rev;               // sweep the tiles last first (odd iterations, large lattices)
    int blk_shift;         // log2 of the tiles per XCD block (0: dispatch order)
    // Q16: h_q = qo[q] + qs[q] * i (decode); i = rint(h_q * qi[q] + qc[q]),
    // clamped to [-32767, 32767] (encode); qi = 1/qs, qc = -qo/qs.  qtab holds
    // the same four rows in device memory for directions known only at run
    // time (obstacle links): indexing these arrays with a run-time value would
    // move the whole parameter block to scratch memory.
    float qs[9], qo[9], qi[9], qc[9];
    const float *qtab;     // [4][9]: qs, qo, qi, qc
};

// ---------------------------------------------------------------- storage --
// A state buffer holds nine population planes of N nodes (plane q, node
// x*ny + y).  ld(p, buf, q, node) returns the population f_q; the float forms
// store a deviation and ldh returns h_q = f_q - w_q; sth stores h_q.
//   F64             f_q in double
//   F32, F16, BF16  h_q in float, IEEE half or bfloat16
//   F16E            g_q = h_q - E_q(v) in IEEE half, with
//                   E_q(v) = w_q (3 c_q.v + 4.5 (c_q.v)^2 - 1.5 v.v)
//                   the equilibrium deviation at the node's velocity v rounded
//                   to half precision; a tenth plane holds v as packed (vx, vy)
//                   half pairs.  g_q is the density and non-equilibrium part
//                   of h_q (~1e-3 against ~2e-2 for h_q in a u ~ 0.05 flow), so
//                   half precision resolves it 10-30 times finer.  Writer and
//                   reader evaluate E_q from the same half-precision v with the
//                   same operations, so the shift cancels exactly.
//   Q16             h_q in 16-bit fixed point on a per-direction affine map,
//                   h_q = qo[q] + qs[q] i with i in [-32767, 32767]: the 65535
//                   levels cover the range of h_q the flow can reach (set by
//                   the caller from its velocity bound), uniformly, where IEEE
//                   half spends most of its codes near zero.  Values outside
//                   the range are clamped to it.
template <int S> struct Store;

__host__ __device__ constexpr bool is16(int S)
{
    return S == LBM_STORE_F16 || S == LBM_STORE_BF16 || S == LBM_STORE_F16E ||
           S == LBM_STORE_Q16 || S == LBM_STORE_F16E8;
}

// Float -> half conversions take their operand through an empty asm, so the
// compiler cannot fold the preceding multiply-add into a mixed-precision FMA
// that rounds to half once (v_fma_mixlo_f16): every kernel then rounds twice,
// to float and to half, and all kernel shapes store identical values.
__device__ __forceinline__ _Float16 to_half(float v)
{
    asm("" : "+v"(v));
    return (_Float16)v;
}

// F16E: offset (in halves) of the velocity plane, 256-byte aligned
__host__ __device__ constexpr size_t vel_offset(int N) { return ((size_t)9 * N + 127) / 128 * 128; }

// 16-bit forms: a plane element through a base pointer that is uniform over
// the work-group (kept in scalar registers) and a 32-bit byte offset shared
// by all planes of a node, so that each access is one global load or store
// with a scalar base and no 64-bit address arithmetic per plane (offsets of
// 2-byte elements stay below 2^32 for any int node count).
template <class E>
__device__ __forceinline__ E ld_off(const void *base, uint32_t off)
{
    return *reinterpret_cast<const E *>(static_cast<const char *>(base) + off);
}
template <class E>
__device__ __forceinline__ void st_off(void *base, uint32_t off, E v)
{
    using gchar = __attribute__((address_space(1))) char;
    using gE = __attribute__((address_space(1))) E;
    *(gE *)((gchar *)base + off) = v;
}
// A uniform pointer through readfirstlane: the compiler keeps each plane's
// base in scalar registers instead of deriving one plane's address from the
// previous one with per-lane 64-bit additions.
template <class T>
__device__ __forceinline__ T *sbase(T *ptr)
{
    const uint64_t v = reinterpret_cast<uint64_t>(ptr);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return reinterpret_cast<T *>(((uint64_t)hi << 32) | lo);
}

template <> struct Store<LBM_STORE_F64> {
    using s_t = double;
    using T = double;
    static __device__ __forceinline__ T ld(const Params &p, const s_t *b, int q, size_t c)
    {
        return b[(size_t)q * p.N + c];
    }
    static __device__ __forceinline__ void st(const Params &, s_t *b, int, size_t i, T v)
    {
        b[i] = v;
    }
};

template <> struct Store<LBM_STORE_F32> {
    using s_t = float;
    using T = float;
    static __device__ __forceinline__ float ldh(const Params &p, const s_t *b, int q, size_t c)
    {
        return b[(size_t)q * p.N + c];
    }
    static __device__ __forceinline__ void sth(const Params &, s_t *b, int, size_t i, float h)
    {
        b[i] = h;
    }
    static __device__ __forceinline__ T ld(const Params &p, const s_t *b, int q, size_t c)
    {
        return weight<float>(q) + ldh(p, b, q, c);
    }
};

template <> struct Store<LBM_STORE_F16> {
    using s_t = _Float16;
    using T = float;
    static __device__ __forceinline__ float dec(const Params &, int, s_t raw) { return (float)raw; }
    static __device__ __forceinline__ s_t enc(const Params &, int, float h) { return to_half(h); }
    static __device__ __forceinline__ float ldh(const Params &p, const s_t *b, int q, size_t c)
    {
        return (float)b[(size_t)q * p.N + c];
    }
    static __device__ __forceinline__ void sth(const Params &, s_t *b, int, size_t i, float h)
    {
        b[i] = to_half(h);
    }
    static __device__ __forceinline__ T ld(const Params &p, const s_t *b, int q, size_t c)
    {
        return weight<float>(q) + ldh(p, b, q, c);
    }
};

__device__ __forceinline__ uint16_t f32_to_bf16(float v)
{
    uint32_t u = __float_as_uint(v);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}

template <> struct Store<LBM_STORE_BF16> {
    using s_t = uint16_t;
    using T = float;
    static __device__ __forceinline__ float dec(const Params &, int, s_t raw)
    {
        return __uint_as_float((uint32_t)raw << 16);
    }
    static __device__ __forceinline__ s_t enc(const Params &, int, float h) { return f32_to_bf16(h); }
    static __device__ __forceinline__ float ldh(const Params &p, const s_t *b, int q, size_t c)
    {
        return __uint_as_float((uint32_t)b[(size_t)q * p.N + c] << 16);
    }
    static __device__ __forceinline__ void sth(const Params &, s_t *b, int, size_t i, float h)
    {
        b[i] = f32_to_bf16(h);
    }
    static __device__ __forceinline__ T ld(const Params &p, const s_t *b, int q, size_t c)
    {
        return weight<float>(q) + ldh(p, b, q, c);
    }
};

// Q16 encode: x = h qi + qc with qc = 1.5 * 2^23 - m (one FMA): the sum lies
// in [2^23, 2^24), where floats are the integers, so x is rounded to
// 1.5 * 2^23 + rint(h / qs) - m, whose low 16 bits are the code i in two's
// complement; x is clamped to the codes -32767 .. 32767.  A non-finite h
// gets the code -32768, which the density sum turns back into a NaN.  The
// multiply-add goes through an empty asm so that every kernel shape rounds
// it the same way.
constexpr float Q16_MAGIC = 12582912.0f;   // 1.5 * 2^23
__device__ __forceinline__ uint16_t q16_encode(float h, float qi, float qc)
{
    float x = __builtin_fmaf(h, qi, qc);
    asm("" : "+v"(x));
    const float c = __builtin_fminf(__builtin_fmaxf(x, Q16_MAGIC - 32767.0f), Q16_MAGIC + 32767.0f);
    return x == x ? (uint16_t)(__float_as_uint(c) & 0xffffu) : (uint16_t)0x8000u;
}

__device__ __forceinline__ float q16_decode(int16_t i, float qs, float qo)
{
    return __builtin_fmaf(qs, (float)i, qo);
}

// Affine-map coefficient `row` (0 qs, 1 qo, 2 qi, 3 qc) of direction q: from
// the kernel arguments when q is a compile-time constant after inlining (the
// unrolled node loops), else from the device table.
__device__ __forceinline__ float q16_coef(const Params &p, int row, int q)
{
    if (__builtin_constant_p(q)) {
        const float *r = row == 0 ? p.qs : row == 1 ? p.qo : row == 2 ? p.qi : p.qc;
        return r[q];
    }
    return p.qtab[row * 9 + q];
}

template <> struct Store<LBM_STORE_Q16> {
    using s_t = int16_t;
    using T = float;
    static __device__ __forceinline__ float dec(const Params &p, int q, s_t raw)
    {
        return q16_decode(raw, q16_coef(p, 0, q), q16_coef(p, 1, q));
    }
    static __device__ __forceinline__ s_t enc(const Params &p, int q, float h)
    {
        return (int16_t)q16_encode(h, q16_coef(p, 2, q), q16_coef(p, 3, q));
    }
    static __device__ __forceinline__ float ldh(const Params &p, const s_t *b, int q, size_t c)
    {
        return q16_decode(b[(size_t)q * p.N + c], q16_coef(p, 0, q), q16_coef(p, 1, q));
    }
    static __device__ __forceinline__ void sth(const Params &p, s_t *b, int q, size_t i, float h)
    {
        b[i] = (int16_t)q16_encode(h, q16_coef(p, 2, q), q16_coef(p, 3, q));
    }
    static __device__ __forceinline__ T ld(const Params &p, const s_t *b, int q, size_t c)
    {
        return weight<float>(q) + ldh(p, b, q, c);
    }
};

// F16E: E_q(v), explicit fmaf only, so every caller rounds identically
// c_q.v is taken as +-vx, +-vy or their sum (no multiplications by the
// integer components, which the compiler may not drop without fast-math), and
// E_q = cu (3 w_q + 4.5 w_q cu) - 1.5 w_q v.v with the products of the weights
// folded into constants (three operations after v.v and cu); q may be a
// run-time value (the link path), the selection then yields the same
// operations.  Writer and reader evaluate it on the same half-precision
// velocity, so the shift cancels exactly whatever its own rounding.
__device__ __forceinline__ float eq_shift(int q, float vx, float vy)
{
    const float vv = __builtin_fmaf(vy, vy, vx * vx);
    const float ax = CX[q] > 0 ? vx : -vx, ay = CY[q] > 0 ? vy : -vy;
    const float cu = CY[q] == 0 ? (CX[q] == 0 ? 0.0f : ax) : (CX[q] == 0 ? ay : ax + ay);
    const float w = weight<float>(q);
    return __builtin_fmaf(cu, __builtin_fmaf(4.5f * w, cu, 3.0f * w), (-1.5f * w) * vv);
}

// The velocity word of the shifted forms.  F16E: (vx, vy) as IEEE halves in
// 32 bits.  F16E8: (vx, vy) as 8-bit fixed point in 16 bits, steps of 2^-9,
// |v| <= 127/512 (clamped; the lattice speed of sound is 0.577 and the flows
// stay far below 0.25): the shift then differs from the node's exact
// equilibrium deviation by at most 3 w_q |c_q| 2^-10 more than with halves,
// still well under the stored residual, for 2 bytes less per node each way.
template <int S> struct Vel;
template <> struct Vel<LBM_STORE_F16E> {
    using w_t = uint32_t;
    static __device__ __forceinline__ w_t pack(float ux, float uy)
    {
        const _Float16 hx = to_half(ux), hy = to_half(uy);
        return (uint32_t)__builtin_bit_cast(uint16_t, hx) |
               ((uint32_t)__builtin_bit_cast(uint16_t, hy) << 16);
    }
    static __device__ __forceinline__ void unpack(w_t v, float &vx, float &vy)
    {
        vx = (float)__builtin_bit_cast(_Float16, (uint16_t)(v & 0xffffu));
        vy = (float)__builtin_bit_cast(_Float16, (uint16_t)(v >> 16));
    }
};
template <> struct Vel<LBM_STORE_F16E8> {
    using w_t = uint16_t;
};

// F16E8 velocity word: a = rint(512 vx), b = rint(512 vy), clamped to
// -127..127, as two's-complement bytes (a low).  Rounding by the magic number
// 1.5 * 2^23 (one FMA, round to nearest even, the integer in the low bits of
// the sum) and a median clamp; a and b come back as integer-valued floats.
__device__ __forceinline__ uint16_t e8_pack(float ux, float uy, float &a, float &b)
{
    constexpr float M = 12582912.0f;   // 1.5 * 2^23
    const float xa = __builtin_amdgcn_fmed3f(__builtin_fmaf(ux, 512.0f, M), M - 127.0f, M + 127.0f);
    const float xb = __builtin_amdgcn_fmed3f(__builtin_fmaf(uy, 512.0f, M), M - 127.0f, M + 127.0f);
    a = xa - M;
    b = xb - M;
    return (uint16_t)((__float_as_uint(xa) & 0xffu) | ((__float_as_uint(xb) & 0xffu) << 8));
}

// the two signed bytes of the word as floats (one byte-select convert each)
__device__ __forceinline__ void e8_unpack(uint32_t w, float &a, float &b)
{
    a = (float)((int32_t)(w << 24) >> 24);
    b = (float)((int32_t)(w << 16) >> 24);
}

// E_q of F16E8 on the integers (a, b) = 512 v: with cu = c_q.(a, b) and
// pv = c_q x (a, b) (the component across c_q), v.v = (cu^2 + pv^2) / |c_q|^2
// / 512^2 and
//   E_q = w_q (3 c.v + 4.5 (c.v)^2 - 1.5 v.v) = cu (K2 + K1 cu) - K3 pv^2,
//   E_0 = -K0 (a^2 + b^2),
// K2 = 3 w / 512, K1 = (4.5 - 1.5 / |c|^2) w / 512^2, K3 = 1.5 w / |c|^2 / 512^2.
// cu and pv are small integers (exact in float); the signs of c fold into
// the operands.  q may be a run-time value (the link path) and then yields
// the same operations on the same constants.
__host__ __device__ constexpr double e8_w(int q) { return q == 0 ? 4.0 / 9.0 : (q <= 4 ? 1.0 / 9.0 : 1.0 / 36.0); }
__host__ __device__ constexpr double e8_c2(int q) { return q <= 4 ? 1.0 : 2.0; }
__device__ __forceinline__ float e8_k0() { return (float)(1.5 * e8_w(0) / 262144.0); }
__device__ __forceinline__ float e8_k1(int q) { return (float)((4.5 - 1.5 / e8_c2(q)) * e8_w(q) / 262144.0); }
__device__ __forceinline__ float e8_k2(int q) { return (float)(3.0 * e8_w(q) / 512.0); }
__device__ __forceinline__ float e8_k3(int q) { return (float)(1.5 / e8_c2(q) * e8_w(q) / 262144.0); }
__device__ __forceinline__ float e8_cu(int q, float a, float b)
{
    const float ax = CX[q] > 0 ? a : -a, by = CY[q] > 0 ? b : -b;
    return CY[q] == 0 ? ax : (CX[q] == 0 ? by : ax + by);
}
__device__ __forceinline__ float e8_pv(int q, float a, float b)
{
    // c x v = cx b - cy a
    const float bx = CX[q] > 0 ? b : -b, ay = CY[q] > 0 ? -a : a;
    return CY[q] == 0 ? bx : (CX[q] == 0 ? ay : bx + ay);
}
// E_q itself: the writer stores half(h - E_q), the reader adds the same E_q
// back, so the shift cancels exactly whatever its own rounding (a reader that
// folded the residual into the multiply-add chain would differ from the
// writer by a float rounding of E, a bias the 200 000-iteration Turek run
// shows: max-Cl error 9.8e-4 against 5.9e-4)
__device__ __forceinline__ float e8_shift(int q, float a, float b)
{
    if (q == 0) {
        const float k = -e8_k0();
        return __builtin_fmaf(a * k, a, (b * k) * b);
    }
    const float cu = e8_cu(q, a, b), pv = e8_pv(q, a, b);
    const float t = __builtin_fmaf(e8_k1(q), cu, e8_k2(q));
    return __builtin_fmaf(cu, t, (pv * -e8_k3(q)) * pv);
}
__device__ __forceinline__ float e8_dec(int q, float a, float b, float r)
{
    float e = e8_shift(q, a, b);
    asm("" : "+v"(e));   // E_q as the writer rounds it, not fused with r
    return e + r;
}

// F16E / F16E8: g_q = h_q - E_q(v) in IEEE half with the velocity word in a
// tenth plane (256-byte aligned)
template <int S> struct StoreShifted {
    using s_t = _Float16;
    using T = float;
    using w_t = typename Vel<S>::w_t;
    static __device__ __forceinline__ const w_t *vel(const s_t *b, int N)
    {
        return reinterpret_cast<const w_t *>(b + vel_offset(N));
    }
    static __device__ __forceinline__ float dec(int q, s_t raw, w_t v)
    {
        if constexpr (S == LBM_STORE_F16E8) {
            float a, b;
            e8_unpack(v, a, b);
            return e8_dec(q, a, b, (float)raw);
        } else {
            float vx, vy;
            Vel<S>::unpack(v, vx, vy);
            return eq_shift(q, vx, vy) + (float)raw;
        }
    }
    static __device__ __forceinline__ float ldh(const Params &p, const s_t *b, int q, size_t c)
    {
        return dec(q, b[(size_t)q * p.N + c], vel(b, p.N)[c]);
    }
    static __device__ __forceinline__ T ld(const Params &p, const s_t *b, int q, size_t c)
    {
        return weight<float>(q) + ldh(p, b, q, c);
    }
};
template <> struct Store<LBM_STORE_F16E> : StoreShifted<LBM_STORE_F16E> {};
template <> struct Store<LBM_STORE_F16E8> : StoreShifted<LBM_STORE_F16E8> {};

__host__ __device__ constexpr bool shifted(int S)
{
    return S == LBM_STORE_F16E || S == LBM_STORE_F16E8;
}

enum { MODE_INIT = 0, MODE_STEP = 1, MODE_EPILOGUE = 2, MODE_DRAG = 3 };

__device__ __forceinline__ int node_x(const Params &p, int cell)
{
    int x = (int)__umulhi((uint32_t)cell, p.magic);
    if (cell - x * p.ny < 0) --x;  // the estimate is never low, at most one high
    return x;
}

__device__ __forceinline__ int bc_row(const Params &p)
{
    int r = (*p.base + p.k) % p.ring;
    return r < 0 ? r + p.ring : r;
}

__device__ __forceinline__ void link_range(const Params &p, int cell, int &lo, int &hi)
{
    lo = hi = 0;
    if (!p.tile_slot) return;
    const int slot = p.tile_slot[cell >> p.tile_shift];
    if (slot < 0) return;
    const int tc = 1 << p.tile_shift;
    const int *off = p.cell_off + (size_t)slot * (tc + 1) + (cell & (tc - 1));
    lo = off[0];
    hi = off[1];
}

template <int S>
__device__ __forceinline__ void pull_checked(const Params &p, int x, int y,
                                             typename Store<S>::T *f)
{
    using St = Store<S>;
    using T = typename St::T;
    const auto *src = static_cast<const typename St::s_t *>(p.src);
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int sx = x - CX[q], sy = y - CY[q];
        f[q] = (sx >= 0 && sx < p.nx && sy >= 0 && sy < p.ny)
                   ? St::ld(p, src, q, (size_t)sx * p.ny + sy)
                   : T(0);
    }
}

// Deviation value h_qb of one link in the float forms.  The two or three
// sources of a link share the weight w_q (planes q and qb are opposite) and
// the interpolation factors sum to one, so f - w obeys the same affine map:
// the rest state bounces back exactly and float-rounded factors cost a
// relative 1e-7 of the deviation, not of w.
template <int S>
__device__ __forceinline__ float link_h(const Params &p, int cell, const LinkRec &L, float &A)
{
    using St = Store<S>;
    const auto *src = static_cast<const typename St::s_t *>(p.src);
    A = St::ldh(p, src, L.q, cell);
    if (!L.mode) return A;
    const float B = St::ldh(p, src, L.planeB, L.cellB);
    const float C = St::ldh(p, src, L.planeC, L.cellC);
    return ((float)L.c0 * A + (float)L.c1 * B) + (float)L.c2 * C;
}

// Obstacle links of a node on full populations f (nb.py:77-117): f[qb] of
// every link of `cell` (links lo..hi-1).
template <int S>
__device__ __forceinline__ void apply_links(const Params &p, int cell, int lo, int hi,
                                            typename Store<S>::T *f)
{
    using St = Store<S>;
    using T = typename St::T;
    const auto *src = static_cast<const typename St::s_t *>(p.src);
    for (int k = lo; k < hi; ++k) {
        const LinkRec L = p.links[k];
        T v;
        if constexpr (S == LBM_STORE_F64) {
            const T A = St::ld(p, src, L.q, cell);
            v = A;
            if (L.mode) {
                const T B = St::ld(p, src, L.planeB, L.cellB);
                const T C = St::ld(p, src, L.planeC, L.cellC);
                v = (T(L.c0) * A + T(L.c1) * B) + T(L.c2) * C;
            }
        } else {
            float A;
            v = weight<float>(L.qb) + link_h<S>(p, cell, L, A);
        }
        // f[L.qb] = v with a register-resident array
#pragma unroll
        for (int q = 1; q < 9; ++q)
            if (q == L.qb) f[q] = v;
    }
}

// The same on deviations h (float forms, interior nodes).
template <int S>
__device__ __forceinline__ void apply_links_h(const Params &p, int cell, int lo, int hi,
                                              float *h)
{
    for (int k = lo; k < hi; ++k) {
        const LinkRec L = p.links[k];
        float A;
        const float v = link_h<S>(p, cell, L, A);
#pragma unroll
        for (int q = 1; q < 9; ++q)
            if (q == L.qb) h[q] = v;
    }
}

// Wall input `side` of iteration row, element i (component-major, as the
// reference's u_left[2][ny] etc.).  Pool form: a table of wall arrays per
// iteration.  Separable form (batched cases): amplitude(row, side) x profile,
// the reference apps' (ret*u_lbm) * profile_j values.
__device__ __forceinline__ double side_value(const Params &p, int row, int side, int i)
{
    if (p.amp) return p.amp[(size_t)row * p.amp_stride + side] * p.prof[side][i];
    return p.pool[p.inlet_off[row * 5 + side] + i];
}

// Post-boundary populations of a node that is not a domain corner; r, ux, uy
// receive what the wall routine writes to rho/u (perimeter nodes only).
template <int S>
__device__ void gather_edge(const Params &p, int x, int y, int cell, int row,
                            typename Store<S>::T *f, typename Store<S>::T &r,
                            typename Store<S>::T &ux, typename Store<S>::T &uy)
{
    using T = typename Store<S>::T;
    pull_checked<S>(p, x, y, f);
    int lo, hi;
    link_range(p, cell, lo, hi);
    apply_links<S>(p, cell, lo, hi, f);
    const int lx = p.nx - 1, ly = p.ny - 1;
    if (y == 0 && p.bc_bottom) {
        ux = T(side_value(p, row, LBM_SIDE_BOTTOM, x));
        uy = T(side_value(p, row, LBM_SIDE_BOTTOM, p.nx + x));
        zh_bottom(f, ux, uy, r);
    } else if (y == ly && p.bc_top) {
        ux = T(side_value(p, row, LBM_SIDE_TOP, x));
        uy = T(side_value(p, row, LBM_SIDE_TOP, p.nx + x));
        zh_top(f, ux, uy, r);
    } else if (x == 0 && p.bc_left) {
        ux = T(side_value(p, row, LBM_SIDE_LEFT, y));
        uy = T(side_value(p, row, LBM_SIDE_LEFT, p.ny + y));
        zh_left(f, ux, uy, r);
    } else if (x == lx && p.bc_right == LBM_BC_VELOCITY) {
        ux = T(side_value(p, row, LBM_SIDE_RIGHT, y));
        uy = T(side_value(p, row, LBM_SIDE_RIGHT, p.ny + y));
        zh_right_vel(f, ux, uy, r);
    } else if (x == lx && p.bc_right == LBM_BC_PRESSURE) {
        r = T(side_value(p, row, LBM_SIDE_RHO_RIGHT, y));
        uy = T(side_value(p, row, LBM_SIDE_RIGHT, p.ny + y));
        zh_right_pres(f, r, uy, ux);
    }
}

// Post-boundary populations of any node (corners recompute the adjacent
// top/bottom wall node, whose rho/u the reference corner routines copy).
template <int S>
__device__ void gather_bc(const Params &p, int x, int y, int cell, int row,
                          typename Store<S>::T *f, typename Store<S>::T &r,
                          typename Store<S>::T &ux, typename Store<S>::T &uy)
{
    using T = typename Store<S>::T;
    const int lx = p.nx - 1, ly = p.ny - 1;
    const bool xe = (x == 0 || x == lx), ye = (y == 0 || y == ly);
    if (!(xe && ye) || !p.corners) {
        gather_edge<S>(p, x, y, cell, row, f, r, ux, uy);
        return;
    }
    // corner: own streamed populations and links, then the corner routine
    pull_checked<S>(p, x, y, f);
    int lo, hi;
    link_range(p, cell, lo, hi);
    apply_links<S>(p, cell, lo, hi, f);
    const int nxn = (x == 0) ? 1 : lx - 1;
    const int which = (x == 0) ? (y == 0 ? 0 : 1) : (y == 0 ? 3 : 2);
    T fn[9];
    gather_edge<S>(p, nxn, y, nxn * p.ny + y, row, fn, r, ux, uy);
    zh_corner(f, which, r, ux, uy);
}

// Drag ring entry of every link of `cell`: f_q(x) + f_qb(x), the pair the
// momentum exchange sums (nb.py:49-73), in double.  The float forms add the
// exact double weights to the deviations, so the weights cancel over an
// obstacle as they do in FP64.
template <int S>
__device__ __forceinline__ void store_drag_range(const Params &p, int cell, int row, int lo,
                                                 int hi, const typename Store<S>::T *f)
{
    using St = Store<S>;
    const auto *src = static_cast<const typename St::s_t *>(p.src);
    for (int k = lo; k < hi; ++k) {
        const LinkRec L = p.links[k];
        double v;
        if constexpr (S == LBM_STORE_F64) {
            double fb = 0.0;
#pragma unroll
            for (int q = 1; q < 9; ++q)
                if (q == L.qb) fb = f[q];
            v = St::ld(p, src, L.q, cell) + fb;
        } else {
            // f[qb] = fl(w + h) with |h| < w/2, so f[qb] - w is exact (Sterbenz)
            float hb = 0.0f;
#pragma unroll
            for (int q = 1; q < 9; ++q)
                if (q == L.qb) hb = f[q] - weight<float>(q);
            const double hA = (double)St::ldh(p, src, L.q, cell);
            v = 2.0 * weight<double>(L.q) + (hA + (double)hb);
        }
        p.drag_ring[(size_t)row * p.n_links + L.dslot] = v;
    }
}

// The same from deviations h (float forms, interior nodes).
template <int S>
__device__ __forceinline__ void store_drag_range_h(const Params &p, int cell, int row, int lo,
                                                   int hi, const float *h)
{
    using St = Store<S>;
    const auto *src = static_cast<const typename St::s_t *>(p.src);
    for (int k = lo; k < hi; ++k) {
        const LinkRec L = p.links[k];
        float hb = 0.0f;
#pragma unroll
        for (int q = 1; q < 9; ++q)
            if (q == L.qb) hb = h[q];
        const double hA = (double)St::ldh(p, src, L.q, cell);
        p.drag_ring[(size_t)row * p.n_links + L.dslot] =
            2.0 * weight<double>(L.q) + (hA + (double)hb);
    }
}

template <int S>
__device__ __forceinline__ void store_drag(const Params &p, int cell, int row,
                                           const typename Store<S>::T *f)
{
    int lo, hi;
    link_range(p, cell, lo, hi);
    store_drag_range<S>(p, cell, row, lo, hi, f);
}

// Moments, equilibrium, TRT collision and store (lattice.py:178-189,
// nb.py:7-17, nb.py:24-35).
template <int S>
__device__ __forceinline__ void collide_store(const Params &p, int cell,
                                              typename Store<S>::T *f, bool macro)
{
    using St = Store<S>;
    using T = typename St::T;
    auto *dst = static_cast<typename St::s_t *>(p.dst);
    T r = f[0];
#pragma unroll
    for (int q = 1; q < 9; ++q) r = r + f[q];
    const T sx = ((((f[1] - f[2]) + f[5]) - f[6]) - f[7]) + f[8];
    const T sy = ((((f[3] - f[4]) + f[5]) - f[6]) + f[7]) - f[8];
    const T ux = sx / r, uy = sy / r;
    if (macro) {
        p.rho_out[cell] = (double)r;
        p.u_out[cell] = (double)ux;
        p.u_out[(size_t)p.N + cell] = (double)uy;
    }
    const T v = T(1.5) * (ux * ux + uy * uy);
    T e[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const T t = T(3.0) * (ux * T(CX[q]) + uy * T(CY[q]));
        e[q] = (T(1.0) + t + T(0.5) * (t * t) - v) * (r * weight<T>(q));
    }
    St::st(p, dst, 0, cell, T(p.b0) * f[0] + T(p.om_p) * e[0]);
    const T a1 = T(p.a1), a2 = T(p.a2), a3 = T(p.a3), a4 = T(p.a4);
#pragma unroll
    for (int q = 1; q < 9; ++q) {
        const int qb = NS[q];
        St::st(p, dst, q, (size_t)q * p.N + cell, a1 * f[q] - a2 * f[qb] + a3 * e[q] + a4 * e[qb]);
    }
}

// Float forms: moments, equilibrium and TRT collision on the deviations
// h_q = f_q - w_q, with the collision written per opposite pair (q, qb):
//   h+- = (h_q +- h_qb)/2,  E+ = w (drho + rho (4.5 cu^2 - 1.5 u^2)),
//   E- = 3 w rho cu,  h_q' = h_q - om_p (h+ - E+) - om_m (h- - E-),
//   h_qb' = h_qb - om_p (h+ - E+) + om_m (h- - E-),
// which is nb.py:24-35 with g_eq of nb.py:7-17, rearranged (E = g_eq - w).
struct Coll {
    float op, hp, hm;  // om_p, om_p/2, om_m/2
};

__device__ __forceinline__ Coll coll_consts(const Params &p)
{
    return Coll{p.omf_p, 0.5f * p.omf_p, 0.5f * p.omf_m};
}

// In place: h (deviations of one node) -> post-collision deviations.
// Every multiply-add is an explicit fmaf and nothing else is contracted (the
// library builds with -ffp-contract=off), so each kernel that inlines this
// -- one node per lane, four nodes per lane, batched -- rounds identically.
__device__ __forceinline__ void collide_h(const Coll &k, float *h, float &rho, float &ux,
                                          float &uy)
{
    const float drho = ((h[0] + h[1]) + (h[2] + h[3])) + ((h[4] + h[5]) + (h[6] + h[7])) + h[8];
    const float jx = (h[1] - h[2]) + (h[5] - h[6]) + (h[8] - h[7]);
    const float jy = (h[3] - h[4]) + (h[5] - h[6]) + (h[7] - h[8]);
    rho = 1.0f + drho;
    const float inv = __builtin_amdgcn_rcpf(rho);
    ux = jx * inv;
    uy = jy * inv;
    const float usq = 1.5f * __builtin_fmaf(uy, uy, ux * ux);
    const float base = __builtin_fmaf(-rho, usq, drho);   // E+/w without the 4.5 cu^2 term
    const float e0 = weight<float>(0) * base;
    h[0] = __builtin_fmaf(-k.op, h[0] - e0, h[0]);
    const float wr3[2] = {3.0f * weight<float>(1) * rho, 3.0f * weight<float>(5) * rho};
    const float w45[2] = {4.5f * weight<float>(1) * rho, 4.5f * weight<float>(5) * rho};
    const float wb[2] = {weight<float>(1) * base, weight<float>(5) * base};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int q = 2 * j + 1, qb = q + 1;   // pairs (1,2) (3,4) (5,6) (7,8)
        const int d = j < 2 ? 0 : 1;
        const float cu = j == 0 ? ux : (j == 1 ? uy : (j == 2 ? ux + uy : uy - ux));
        const float ep = __builtin_fmaf(w45[d] * cu, cu, wb[d]);
        const float em = wr3[d] * cu;
        const float dp = k.hp * __builtin_fmaf(-2.0f, ep, h[q] + h[qb]);
        const float dm = k.hm * __builtin_fmaf(-2.0f, em, h[q] - h[qb]);
        const float a = (h[q] - dp) - dm;
        const float bb = (h[qb] - dp) + dm;
        h[q] = a;
        h[qb] = bb;
    }
}

template <int S>
__device__ __forceinline__ void collide_store_h(const Params &p, int cell, const float *hin,
                                                bool macro)
{
    using St = Store<S>;
    auto *dst = static_cast<typename St::s_t *>(p.dst);
    float h[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) h[q] = hin[q];
    float rho, ux, uy;
    collide_h(coll_consts(p), h, rho, ux, uy);
    if (macro) {
        p.rho_out[cell] = (double)rho;
        p.u_out[cell] = (double)ux;
        p.u_out[(size_t)p.N + cell] = (double)uy;
    }
    if constexpr (shifted(S)) {
        // the node's velocity word, then g_q = h_q - E_q(v) of the stored v
        using w_t = typename Vel<S>::w_t;
        const uint32_t off = (uint32_t)cell * (uint32_t)sizeof(typename St::s_t);
        if constexpr (S == LBM_STORE_F16E8) {
            float a, b;
            st_off(sbase(dst + vel_offset(p.N)), off, e8_pack(ux, uy, a, b));
#pragma unroll
            for (int q = 0; q < 9; ++q)
                st_off(sbase(dst + (size_t)q * p.N), off, to_half(h[q] - e8_shift(q, a, b)));
        } else {
            const w_t v = Vel<S>::pack(ux, uy);
            reinterpret_cast<w_t *>(dst + vel_offset(p.N))[cell] = v;
            float vx, vy;
            Vel<S>::unpack(v, vx, vy);
#pragma unroll
            for (int q = 0; q < 9; ++q)
                st_off(sbase(dst + (size_t)q * p.N), off, to_half(h[q] - eq_shift(q, vx, vy)));
        }
    } else if constexpr (is16(S)) {
        const uint32_t off = (uint32_t)cell * (uint32_t)sizeof(typename St::s_t);
#pragma unroll
        for (int q = 0; q < 9; ++q) st_off(sbase(dst + (size_t)q * p.N), off, St::enc(p, q, h[q]));
    } else {
#pragma unroll
        for (int q = 0; q < 9; ++q) St::sth(p, dst, q, (size_t)q * p.N + cell, h[q]);
    }
}

// Work-group index -> tile index, so that each XCD (work-groups are dealt
// round-robin over the 8 XCDs) sweeps one contiguous slab of the lattice and
// the rows its tiles share stay in that XCD's L2 (the initial and final
// passes; the step sweeps tiles in dispatch order, sweep_tile).
// With rev set the slab is swept last tile first.
__device__ __forceinline__ int xcd_tile(int b, int nb, int rev = 0)
{
    const int q = nb >> 3, rr = nb & 7;
    const int xcd = b & 7, i = b >> 3;
    const int len = q + (xcd < rr ? 1 : 0);   // work-groups of this XCD
    return xcd * q + min(xcd, rr) + (rev ? len - 1 - i : i);
}

template <int S>
__device__ __forceinline__ void collide_any(const Params &p, int cell,
                                            typename Store<S>::T *f, bool macro)
{
    if constexpr (S == LBM_STORE_F64) {
        collide_store<S>(p, cell, f, macro);
    } else {
        float h[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) h[q] = f[q] - weight<float>(q);
        collide_store_h<S>(p, cell, h, macro);
    }
}

// Initial and final passes, one node per work-item (grid = ceil(N/256/CPT)).
template <int S, int MODE, int CPT>
__global__ __launch_bounds__(256) void k_fused(Params p)
{
    using St = Store<S>;
    using T = typename St::T;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int base = tile * (256 * CPT) + threadIdx.x;
    const int lx = p.nx - 1, ly = p.ny - 1;
    const int row = (MODE == MODE_INIT) ? 0 : bc_row(p);
    const bool macro = p.rho_out != nullptr;
    const auto *src = static_cast<const typename St::s_t *>(p.src);
#pragma unroll 1
    for (int c = 0; c < CPT; ++c) {
        const int cell = base + c * 256;
        if (cell >= p.N) return;
        T f[9];
        if (MODE == MODE_INIT) {
            // populations of iteration 0; without an array, the rest state g = w
#pragma unroll
            for (int q = 0; q < 9; ++q)
                f[q] = p.g0 ? T(p.g0[(size_t)q * p.N + cell]) : weight<T>(q);
            collide_any<S>(p, cell, f, macro);
            continue;
        }
        const int x = node_x(p, cell), y = cell - x * p.ny;
        const bool interior = (unsigned)(x - 1) < (unsigned)(lx - 1) &&
                              (unsigned)(y - 1) < (unsigned)(ly - 1);
        T r = T(0), ux = T(0), uy = T(0);
        gather_bc<S>(p, x, y, cell, row, f, r, ux, uy);
        store_drag<S>(p, cell, row, f);
        if (MODE == MODE_STEP) {
            collide_any<S>(p, cell, f, macro);
        } else {  // MODE_EPILOGUE
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                p.g_out[(size_t)q * p.N + cell] = (double)f[q];
                // the float forms' stored populations are w_q + h_q with the
                // exact double weight w_q
                if constexpr (S == LBM_STORE_F64)
                    p.gup_out[(size_t)q * p.N + cell] = St::ld(p, src, q, cell);
                else
                    p.gup_out[(size_t)q * p.N + cell] =
                        weight<double>(q) + (double)St::ldh(p, src, q, cell);
            }
            if (!interior && p.rho_out) {
                p.rho_out[cell] = (double)r;
                p.u_out[cell] = (double)ux;
                p.u_out[(size_t)p.N + cell] = (double)uy;
            }
        }
    }
}

// Perimeter node i of the perimeter work-groups -> (x, y).  Every perimeter
// node (n_perim = 2nx + 2ny - 4): the bottom row, the top row, then the left
// and right columns without their corners.  COLS (n_perim = 2ny, the tiles
// complete the bottom and top walls): the left column, then the right one,
// corners included.
template <bool COLS>
__device__ __forceinline__ void perim_node(const Params &p, int i, int &x, int &y)
{
    const int nx = p.nx, ny = p.ny;
    if (COLS && p.n_perim == 2 * ny) {   // FP64/FP32; 16-bit forms and f16e if ny >= 64
        x = i < ny ? 0 : nx - 1;
        y = i < ny ? i : i - ny;
        return;
    }
    if (i < nx) { x = i; y = 0; return; }
    i -= nx;
    if (i < nx) { x = i; y = ny - 1; return; }
    i -= nx;
    if (i < ny - 2) { x = 0; y = i + 1; return; }
    i -= ny - 2;
    x = nx - 1;
    y = i + 1;
}

// Perimeter work-groups: the first n_perim perimeter nodes in perim_node
// order, node by node through the Zou-He plan.
template <int S, bool COLS>
__device__ __forceinline__ void step_perimeter(const Params &p, int pblock, int row, bool macro)
{
    using T = typename Store<S>::T;
    const int i = pblock * 256 + threadIdx.x;
    if (i >= p.n_perim) return;
    int x, y;
    perim_node<COLS>(p, i, x, y);
    const int cell = x * p.ny + y;
    T f[9], r = T(0), ux = T(0), uy = T(0);
    gather_bc<S>(p, x, y, cell, row, f, r, ux, uy);
    store_drag<S>(p, cell, row, f);
    collide_any<S>(p, cell, f, macro);
}

// The one-node tiles complete their bottom and top wall nodes whenever the
// perimeter work-groups take the left and right columns only (n_perim = 2ny:
// every form except with the four-node kernel; the 16-bit forms and f16e need
// ny >= 64, see wall_routine).  On the 8192x2048 obstacle array this took FP64
// from 491 to 460 us, FP32 from 256 to 240 us, FP16 from 138 to 133 us and
// f16e from 206 to 199 us per iteration: the wall nodes' strided accesses in
// the perimeter work-groups become part of the tile's coalesced ones.

// Bottom or top wall routine of a tile's wall node (not a corner) on the
// populations its work-item pulled: hv holds f (FP64) or h = f - w (float
// forms), where f = w + h is exactly Store::ld.  The float forms go through f
// and back as the perimeter path does (collide_any), so a wall node completed
// in a tile rounds exactly as one completed by a perimeter work-group.  The
// 16-bit forms and f16e (ny >= 64: a wave's run of 64 consecutive nodes
// holds at most one bottom and one top wall node, whose rows follow from the
// run's first node) read the wall velocity through the scalar cache and
// evaluate both wall routines, selecting the result: this keeps every
// population in registers and the register peak low (FP16: 77 VGPRs).
template <int S>
__device__ __forceinline__ void wall_routine(const Params &p, int cell, int row, bool bottom,
                                             typename Store<S>::T *hv)
{
    using T = typename Store<S>::T;
    const int x = node_x(p, cell);
    if constexpr (is16(S)) {
        // wall velocities are read-only during the launch: scalar loads
        using cd = const __attribute__((address_space(4))) double;
        using ci = const __attribute__((address_space(4))) int;
        const int first = __builtin_amdgcn_readfirstlane(cell - (int)(threadIdx.x & 63));
        const int xs = bottom ? (first + p.ny - 1) / p.ny : (first + 63 - (p.ny - 1)) / p.ny;
        const int xc = min(max(xs, 0), p.nx - 1);
        const int side = bottom ? LBM_SIDE_BOTTOM : LBM_SIDE_TOP;
        double dux, duy;
        if (p.amp) {
            const double a = ((cd *)p.amp)[(size_t)row * p.amp_stride + side];
            dux = a * ((cd *)p.prof[side])[xc];
            duy = a * ((cd *)p.prof[side])[p.nx + xc];
        } else {
            const int off = ((ci *)p.inlet_off)[row * 5 + side];
            dux = ((cd *)p.pool)[off + xc];
            duy = ((cd *)p.pool)[off + p.nx + xc];
        }
        // f-space Zou-He exactly as the perimeter path (zh_bottom / zh_top on
        // f = w + h, then h = f - w), both evaluated and selected so that no
        // store goes through a selected element
        const bool apply = bottom ? p.bc_bottom : p.bc_top;
        const float ux = (float)dux, uy = (float)duy;
        float f0[9], fb[9], ft[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) f0[q] = fb[q] = ft[q] = weight<float>(q) + hv[q];
        float rb, rt;
        zh_bottom(fb, ux, uy, rb);
        zh_top(ft, ux, uy, rt);
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const float fn = apply ? (bottom ? fb[q] : ft[q]) : f0[q];
            hv[q] = fn - weight<float>(q);
        }
        return;
    }
    T f[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        if constexpr (S == LBM_STORE_F64)
            f[q] = hv[q];
        else
            f[q] = weight<float>(q) + hv[q];
    }
    T r = T(0), ux = T(0), uy = T(0);
    if (bottom && p.bc_bottom) {
        ux = T(side_value(p, row, LBM_SIDE_BOTTOM, x));
        uy = T(side_value(p, row, LBM_SIDE_BOTTOM, p.nx + x));
        zh_bottom(f, ux, uy, r);
    } else if (!bottom && p.bc_top) {
        ux = T(side_value(p, row, LBM_SIDE_TOP, x));
        uy = T(side_value(p, row, LBM_SIDE_TOP, p.nx + x));
        zh_top(f, ux, uy, r);
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        if constexpr (S == LBM_STORE_F64)
            hv[q] = f[q];
        else
            hv[q] = f[q] - weight<float>(q);
    }
}

// Obstacle links of one wave's run of 64 nodes, link-major (float forms):
// work-item k takes record r0 + k (r0 + k + 64, ... for runs with more links),
// loads it, reads the link's sources and writes the link value to the owning
// node's slot in the wave's LDS area (lv[node][qb - 1], flag bit qb in
// lf[node]), and the drag term of an interior owner to the ring.  A node with
// several links thus costs two dependent loads (record, sources) instead of
// two per link.  Same values as apply_links_h / store_drag_range_h.  All 64
// work-items of the wave take part (the caller branches on the wave-uniform
// range); LDS operations of a wave complete in order.
template <int S>
__device__ __forceinline__ void link_pass(const Params &p, int r0, int r1, int first, int row,
                                          float *lv, uint32_t *lf)
{
    const int lane = threadIdx.x & 63;
    const int lx = p.nx - 1, ly = p.ny - 1;
    lf[lane] = 0u;
    for (int k0 = r0; k0 < r1; k0 += 64) {
        const int k = k0 + lane;
        if (k < r1) {
            const LinkRec L = p.links[k];
            float A;   // h_q of the link's node: the first source and the drag term's
            const float v = link_h<S>(p, L.cell, L, A);
            const int o = L.cell - first;
            lv[o * 8 + (L.qb - 1)] = v;
            atomicOr(&lf[o], 1u << L.qb);
            const int x = node_x(p, L.cell), y = L.cell - x * p.ny;
            if ((unsigned)(x - 1) < (unsigned)(lx - 1) && (unsigned)(y - 1) < (unsigned)(ly - 1))
                p.drag_ring[(size_t)row * p.n_links + L.dslot] =
                    2.0 * weight<double>(L.q) + ((double)A + (double)v);
        }
    }
}

// Tile of 256*CPT nodes: every work-item issues the 9*CPT pull loads of its
// nodes first, then fixes up obstacle-link nodes (tiles with links only),
// collides and stores.  Bottom and top wall nodes (corners excepted) are
// completed in the tile with the same loads and stores (wall_routine); the
// left and right columns are left to the perimeter work-groups (their nodes
// load a valid interior node).
template <int S, int CPT>
__device__ __forceinline__ void step_tile(const Params &p, int tile, bool macro)
{
    using St = Store<S>;
    using T = typename St::T;
    using s_t = typename St::s_t;
    const s_t *__restrict__ src = static_cast<const s_t *>(p.src);
    // link tables are read-only during the launch: the tile's slot and each
    // wave's record range come through the scalar cache, after the pull loads
    using ci32 = const __attribute__((address_space(4))) int;
    const int base = tile * (256 * CPT) + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int lx = p.nx - 1, ly = p.ny - 1;
    T v[CPT][9];
    bool in[CPT], wall[CPT], bottom[CPT];
    int rbeg[CPT], rend[CPT];   // float forms: link records of each wave's run
    __shared__ float s_lv[4][64 * 8];   // link values per wave, node, direction
    __shared__ uint32_t s_lf[4][64];    // direction bits of the links received
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const int cell = base + c * 256;
        // a wave's 64 nodes lie in one column when ny is a multiple of 64:
        // its column index then comes from the scalar unit
        int x;
        if ((p.ny & 63) == 0)
            x = node_x(p, __builtin_amdgcn_readfirstlane(cell - lane));
        else
            x = node_x(p, cell);
        const int y = cell - x * p.ny;
        in[c] = cell < p.N && (unsigned)(x - 1) < (unsigned)(lx - 1) &&
                (unsigned)(y - 1) < (unsigned)(ly - 1);
        wall[c] = p.n_perim == 2 * p.ny && cell < p.N &&
                  (unsigned)(x - 1) < (unsigned)(lx - 1) && (y == 0 || y == ly);
        bottom[c] = y == 0;
        const int lc = (in[c] || wall[c]) ? cell : p.ny + 1;
        if constexpr (is16(S)) {
            // plane q shifted by the source offset as a scalar base, the
            // node's byte offset shared by all planes.  A wall node's sources
            // beyond the wall are read from the neighbouring column (index -1
            // or N at the corners of the lattice: the neighbouring plane, or
            // the state buffer's tail padding) and replaced by the wall
            // routine.
            const uint32_t off = (uint32_t)lc * (uint32_t)sizeof(s_t);
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                const ptrdiff_t sh = (ptrdiff_t)CX[q] * p.ny + CY[q];
                const s_t raw = ld_off<s_t>(src + (size_t)q * p.N - sh, off);
                if constexpr (shifted(S)) {
                    using w_t = typename Vel<S>::w_t;
                    const w_t *vb = St::vel(src, p.N) - sh;
                    w_t w;
                    if constexpr (sizeof(w_t) == sizeof(s_t))
                        w = ld_off<w_t>(vb, off);
                    else
                        w = vb[(size_t)lc];
                    v[c][q] = St::dec(q, raw, w);
                } else {
                    v[c][q] = St::dec(p, q, raw);
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                // source node; a wall node's sources outside the lattice read
                // the node itself (the value is replaced by the wall routine)
                int sc = lc - (CX[q] * p.ny + CY[q]);
                if (wall[c] && ((CY[q] == 1 && y == 0) || (CY[q] == -1 && y == ly))) sc = lc;
                if constexpr (S == LBM_STORE_F64)
                    v[c][q] = St::ld(p, src, q, (size_t)sc);
                else
                    v[c][q] = St::ldh(p, src, q, (size_t)sc);
            }
        }
    }
    // the link tables and the iteration's ring row (a device counter) only
    // after the pull loads are issued: their scalar loads then overlap the
    // pulls' latency instead of delaying them.  The kernel arguments of this
    // phase come in one batch, then the tile's word of run bits and the ring
    // counter together: two rounds of scalar loads, and a third only for the
    // runs with links (their record bounds)
    __builtin_amdgcn_sched_barrier(0);
    const int slot = (S == LBM_STORE_F64 && p.tile_slot) ? ((const ci32 *)p.tile_slot)[tile] : -1;
    const int *wrp = p.wave_rng;
    const uint32_t *rbp = p.run_bits;
    const int *basep = p.base;
    const int ring = p.ring, kk = p.k, rstride = p.rng_stride;
    asm volatile("" ::"s"(wrp), "s"(rbp), "s"(basep), "s"(ring), "s"(kk), "s"(rstride));
    // the tile's runs [4 CPT tile, 4 CPT tile + 4 CPT) lie in one word
    const uint32_t r0 = (uint32_t)tile * (4 * CPT);
    uint32_t word = 0;
    if (S != LBM_STORE_F64 && wrp)
        word = ((const __attribute__((address_space(4))) uint32_t *)rbp)[r0 >> 5];
    int row = *(const ci32 *)basep;
    asm volatile("" : "+s"(word), "+s"(row));
    row = (row + kk) % ring;
    row = row < 0 ? row + ring : row;
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        rbeg[c] = rend[c] = 0;
        // run j = c*4 + wave of the tile (nodes [64 j, 64 j + 64)): for a
        // run with links, its records [wave_rng[j], wave_rng[j + 1]) of the
        // tile's row
        const int j = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6) + c * 4);
        if (S != LBM_STORE_F64 && ((word >> ((r0 & 31) + j)) & 1u)) {
            const ci32 *wr = (const ci32 *)wrp + (size_t)tile * rstride + j;
            rbeg[c] = wr[0];
            rend[c] = wr[1];
        }
    }
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        // float forms: the run's links, link-major, before the nodes branch
        uint32_t lflags = 0;
        if constexpr (S != LBM_STORE_F64) {
            if (rend[c] > rbeg[c]) {
                const int w = threadIdx.x >> 6;
                link_pass<S>(p, rbeg[c], rend[c],
                             __builtin_amdgcn_readfirstlane(base + c * 256 - lane), row,
                             s_lv[w], s_lf[w]);
                lflags = s_lf[w][lane];
            }
        }
        if (!in[c] && !wall[c]) continue;
        const int cell = base + c * 256;
        if (wall[c]) {
            // sources outside the lattice are zero (f = 0, h = -w) before the
            // links and the wall routine, as in the perimeter path
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                if ((bottom[c] && CY[q] == 1) || (!bottom[c] && CY[q] == -1)) {
                    if constexpr (S == LBM_STORE_F64)
                        v[c][q] = 0.0;
                    else
                        v[c][q] = -weight<float>(q);
                }
            }
        }
        int lo = 0, hi = 0;
        if constexpr (S == LBM_STORE_F64) {
            // FP64: the node's mask word per work-item (measured faster here
            // than the wave's scalar mask: 461 against 479 us)
            if (slot >= 0) {
                const int local = cell & ((256 * CPT) - 1);
                const uint32_t word = p.cell_mask[(size_t)slot * (8 * CPT) + (local >> 5)];
                if ((word >> (local & 31)) & 1u) {
                    link_range(p, cell, lo, hi);
                    apply_links<S>(p, cell, lo, hi, v[c]);
                }
            }
        } else if (lflags) {
            const float *mine = s_lv[threadIdx.x >> 6] + lane * 8;
#pragma unroll
            for (int q = 1; q < 9; ++q)
                if ((lflags >> q) & 1u) v[c][q] = mine[q - 1];
            // a wall node's drag terms follow its wall routine (below)
            if (wall[c]) link_range(p, cell, lo, hi);
        }
        if (wall[c]) wall_routine<S>(p, cell, row, bottom[c], v[c]);
        if (lo < hi) {
            if constexpr (S == LBM_STORE_F64)
                store_drag_range<S>(p, cell, row, lo, hi, v[c]);
            else
                store_drag_range_h<S>(p, cell, row, lo, hi, v[c]);
        }
        if constexpr (S == LBM_STORE_F64)
            collide_store<S>(p, cell, v[c], macro);
        else
            collide_store_h<S>(p, cell, v[c], macro);
    }
}

// The time step.  Work-groups [0, n_main) sweep interior tiles of 256*CPT
// nodes: every work-item issues the 9*CPT pull loads of its nodes first
// (perimeter nodes of the tile load a valid interior node and are skipped),
// then fixes up obstacle-link nodes (tiles with links only) and collides.
// Work-groups [n_main, grid) apply the Zou-He perimeter plan node by node.
// Work-groups of the one-node step: first the perimeter work-groups (rounded
// up to a multiple of 8, so that tile t runs on the XCD of work-group t), so
// that their node-by-node chains run beside the tiles instead of after them
// (FP16 on the 8192x2048 obstacle array: 127.9 -> 125.2 us), then the tiles.
// Work-group index -> tile of the one-node step: consecutive tiles in
// dispatch order (consecutive work-groups go to different XCDs), the last
// tile first on odd iterations when rev is set, so that each sweep starts on
// the data the previous one wrote last.  Against one contiguous slab per XCD
// (xcd_tile): the eight XCDs then stream neighbouring regions of each plane
// at any time (a data-movement-only kernel of the FP16 pattern on the
// 8192x2048 lattice: 107.4 against 117.5 us).  With bs > 0, blocks of 2^bs
// consecutive tiles (4 on lattices up to 128 MB; one lattice column for the
// float forms above) go to one XCD each, block j to XCD j % 8, and the XCDs
// still sweep neighbouring blocks.
__device__ __forceinline__ int sweep_tile(int b, int nb, int rev, int bs = 0)
{
    const int t = xcd_block_unit(b, nb, bs);
    return rev ? nb - 1 - t : t;
}

__host__ __device__ constexpr int perim_groups(int n_perim)
{
    return (((n_perim + 255) / 256) + 7) & ~7;
}

// One word of every 64-byte line of a parameter block, loaded in one batch:
// the fields a kernel reads are otherwise fetched in several dependent rounds
// of scalar loads, each of which can miss the scalar cache.  The batched step
// reads each work item's block from device memory and gains from it (256
// cavities, FP32: 62.8 -> 65.3 GLUPS); the one-lattice step reads its block
// from the kernel arguments and does not (f16e8 on the 8192x2048 obstacle
// array: 143 -> 150 us with it).
__device__ __forceinline__ void touch_params(const Params &p)
{
    const uint32_t *w = reinterpret_cast<const uint32_t *>(&p);
    constexpr int lines = (int)((sizeof(Params) + 63) / 64);
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < lines; ++i) acc ^= w[i * 16];
    asm volatile("" ::"s"(acc));
}

template <int S, int CPT>
__device__ __forceinline__ void step_launch(const Params &p)
{
    const bool macro = p.rho_out != nullptr;
    const int np = perim_groups(p.n_perim);
    if ((int)blockIdx.x < np)
        step_perimeter<S, true>(p, (int)blockIdx.x, bc_row(p), macro);
    else
        step_tile<S, CPT>(p, sweep_tile((int)blockIdx.x - np, p.n_main, p.rev, p.blk_shift), macro);
}

template <int S, int CPT>
__global__ __launch_bounds__(256) void k_step(Params p)
{
    step_launch<S, CPT>(p);
}

// The same with an occupancy floor of 7 waves per SIMD (72 VGPRs), for the
// shapes whose tile register peak -- a wall node's routine with the other
// nodes' populations live -- only just exceeds it and fits without spilling
// on the node path (FP64 with two nodes per work-item; FP32, FP16 and BF16
// with four: FP16 125.3 -> 125.0 us on the 8192x2048 obstacle array).
template <int S, int CPT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void k_step_w7(Params p)
{
    step_launch<S, CPT>(p);
}

// And with 8 waves per SIMD (64 VGPRs): f16e8 with one node per work-item
// fits without spilling (66 -> 64 VGPRs; 171.6 -> 163.8 us on the 8192x2048
// obstacle array); f16e with one spilled register (65 -> 64 VGPRs; 154.4 ->
// 143.9 us).
template <int S, int CPT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_step_w8(Params p)
{
    step_launch<S, CPT>(p);
}

// 16-bit storage forms, four consecutive nodes per work-item: every plane is
// read and written with aligned 8-byte accesses.  A plane shifted by one node
// in y (cy = +-1) takes its missing element from the neighbouring work-item
// (DPP whole-wave shift) or, at the edges of the wave's 256-node segment, from
// one extra load.  Needs ny % 4 == 0 and N % (1024*CPT) == 0.
__device__ __forceinline__ float h16_to_f(uint32_t b, bool bf16)
{
    if (bf16) return __uint_as_float(b << 16);
    return (float)__builtin_bit_cast(_Float16, (uint16_t)b);
}

// DPP controls of the whole-wave shifts (GFX9 encoding)
constexpr int DPP_WAVE_SHL1 = 0x130;
constexpr int DPP_WAVE_SHR1 = 0x138;

__device__ __forceinline__ uint32_t f_to_h16(float v, bool bf16)
{
    if (bf16) return f32_to_bf16(v);
    return (uint32_t)__builtin_bit_cast(uint16_t, to_half(v));
}

template <int S, int CPT>
__device__ __forceinline__ void step_tile_v4(const Params &p, int tile, int row, bool macro)
{
    constexpr bool BF = (S == LBM_STORE_BF16);
    const uint16_t *__restrict__ src = static_cast<const uint16_t *>(p.src);
    uint16_t *__restrict__ dst = static_cast<uint16_t *>(p.dst);
    const int slot = p.tile_slot ? p.tile_slot[tile] : -1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lx = p.nx - 1, ly = p.ny - 1;
    const Coll kc = coll_consts(p);
#pragma unroll 1
    for (int c = 0; c < CPT; ++c) {
        const int seg = tile * (1024 * CPT) + c * 1024 + wave * 256;  // wave's first node
        const int cell0 = seg + lane * 4;
        // all loads first: nine 8-byte plane loads and, for the six planes
        // shifted in y, the one element beyond the wave's segment
        uint2 raw[9];
        uint32_t edge[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const uint16_t *pq = src + ((ptrdiff_t)q * p.N - CX[q] * p.ny);
            raw[q] = *reinterpret_cast<const uint2 *>(pq + cell0);
            if (CY[q] == 1) edge[q] = pq[seg - 1];
            else if (CY[q] == -1) edge[q] = pq[seg + 256];
        }
        float h[4][9];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const uint2 a = raw[q];
            uint32_t e[4] = {a.x & 0xffffu, a.x >> 16, a.y & 0xffffu, a.y >> 16};
            if (CY[q] == 1) {          // source y-1: (prev, e0, e1, e2)
                // wave_shr:1 -- lane i takes e[3] of lane i-1; lane 0 has no
                // source lane and keeps the old operand, the edge element
                const uint32_t prev = (uint32_t)__builtin_amdgcn_update_dpp(
                    (int)edge[q], (int)e[3], DPP_WAVE_SHR1, 0xf, 0xf, false);
                e[3] = e[2]; e[2] = e[1]; e[1] = e[0]; e[0] = prev;
            } else if (CY[q] == -1) {  // source y+1: (e1, e2, e3, next)
                // wave_shl:1 -- lane i takes e[0] of lane i+1; lane 63 keeps
                // the edge element
                const uint32_t next = (uint32_t)__builtin_amdgcn_update_dpp(
                    (int)edge[q], (int)e[0], DPP_WAVE_SHL1, 0xf, 0xf, false);
                e[0] = e[1]; e[1] = e[2]; e[2] = e[3]; e[3] = next;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) h[k][q] = h16_to_f(e[k], BF);
        }
        const int x = node_x(p, cell0), y0 = cell0 - x * p.ny;
        const bool xin = (unsigned)(x - 1) < (unsigned)(lx - 1);
        bool in[4];
        uint32_t out[4][9];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int cell = cell0 + k;
            in[k] = xin && (unsigned)(y0 + k - 1) < (unsigned)(ly - 1);
            if (in[k] && slot >= 0) {
                const int local = cell & (1024 * CPT - 1);
                const uint32_t word = p.cell_mask[(size_t)slot * (32 * CPT) + (local >> 5)];
                if ((word >> (local & 31)) & 1u) {
                    int lo, hi;
                    link_range(p, cell, lo, hi);
                    apply_links_h<S>(p, cell, lo, hi, h[k]);
                    store_drag_range_h<S>(p, cell, row, lo, hi, h[k]);
                }
            }
            float rho, ux, uy;
            collide_h(kc, h[k], rho, ux, uy);
            if (macro && in[k]) {
                p.rho_out[cell] = (double)rho;
                p.u_out[cell] = (double)ux;
                p.u_out[(size_t)p.N + cell] = (double)uy;
            }
#pragma unroll
            for (int q = 0; q < 9; ++q) out[k][q] = f_to_h16(h[k][q], BF);
        }
        const bool all_in = in[0] && in[1] && in[2] && in[3];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            uint16_t *dq = dst + (size_t)q * p.N;
            if (all_in) {
                uint2 v;
                v.x = out[0][q] | (out[1][q] << 16);
                v.y = out[2][q] | (out[3][q] << 16);
                *reinterpret_cast<uint2 *>(dq + cell0) = v;
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (in[k]) dq[cell0 + k] = (uint16_t)out[k][q];
            }
        }
    }
}

template <int S, int CPT>
__global__ __launch_bounds__(256) void k_step_v4(Params p)
{
    const int row = bc_row(p);
    const bool macro = p.rho_out != nullptr;
    if ((int)blockIdx.x >= p.n_main)
        step_perimeter<S, false>(p, (int)blockIdx.x - p.n_main, row, macro);
    else
        step_tile_v4<S, CPT>(p, sweep_tile(blockIdx.x, p.n_main, p.rev), row, macro);
}

// Batched time step: many lattices in one launch.  cases[2*c + parity] holds
// the parameters of case c for iterations of that parity (ping-pong buffers);
// work[] lists (case, block) items -- the interior tiles and perimeter blocks
// of every case -- swept grid-stride by a grid sized to the GPU.
template <int S, int CPT, bool V4>
__global__ __launch_bounds__(256) void k_batch_step(const Params *__restrict__ cases,
                                                    const int2 *__restrict__ work, int n_work,
                                                    int k, const int *base, int parity)
{
    for (int w = blockIdx.x; w < n_work; w += gridDim.x) {
        const int2 item = work[w];
        touch_params(cases[2 * item.x + parity]);
        Params p = cases[2 * item.x + parity];
        p.k = k;
        p.base = base;
        if (item.y >= p.n_main) {
            step_perimeter<S, !V4>(p, item.y - p.n_main, bc_row(p), false);
        } else {
            if constexpr (V4)
                step_tile_v4<S, CPT>(p, item.y, bc_row(p), false);
            else
                step_tile<S, CPT>(p, item.y, false);
        }
    }
}

// Drag terms of one iteration on the link nodes only.
template <int S>
__global__ __launch_bounds__(256) void k_drag_only(Params p)
{
    using T = typename Store<S>::T;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= p.n_link_cells) return;
    const int cell = p.link_cells[i];
    const int x = node_x(p, cell), y = cell - x * p.ny;
    const int row = bc_row(p);
    T f[9], r = T(0), ux = T(0), uy = T(0);
    gather_bc<S>(p, x, y, cell, row, f, r, ux, uy);
    store_drag<S>(p, cell, row, f);
}

// Per-obstacle sums of the ring rows, in link plan order (nb.py:57-67).
__global__ void k_drag_reduce(const double *__restrict__ ringbuf, int n_links, int ring,
                              int64_t it0, int n, const int *__restrict__ obs_off, int n_obs,
                              const int8_t *__restrict__ plan_q, double *__restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * n_obs) return;
    const int s = t / n_obs, o = t - s * n_obs;
    const int row = (int)((it0 + s) % ring);
    const double *r = ringbuf + (size_t)row * n_links;
    double fx = 0.0, fy = 0.0;
    for (int k = obs_off[o]; k < obs_off[o + 1]; ++k) {
        const double g0 = r[k];
        fx = fx + g0 * (double)CX[plan_q[k]];
        fy = fy + g0 * (double)CY[plan_q[k]];
    }
    out[2 * (size_t)t] = fx;
    out[2 * (size_t)t + 1] = fy;
}

// Density sum of one state buffer in double: FP64 sums f_q; the float forms
// sum their stored deviations h_q in double (their populations are
// w_q + h_q with the exact double weights w_q), and k_sum_serial adds
// N * sum_q w_q once.  Work-group partials land in part[blockIdx.x].
template <int S>
__global__ __launch_bounds__(256) void k_mass_partial(Params p, double *part)
{
    using St = Store<S>;
    const auto *src = static_cast<const typename St::s_t *>(p.src);
    const int N = p.N;
    __shared__ double sh[256];
    double acc = 0.0;
    for (int c = blockIdx.x * 256 + threadIdx.x; c < N; c += gridDim.x * 256) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            if constexpr (S == LBM_STORE_F64)
                s += St::ld(p, src, q, c);
            else
                s += (double)St::ldh(p, src, q, c);
            if constexpr (S == LBM_STORE_Q16)   // the code of a non-finite value
                if (src[(size_t)q * N + c] == (int16_t)-32768) s = __builtin_nan("");
        }
        acc += s;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// out = sum of the n partials + base (the weights' share of the float forms)
__global__ void k_sum_serial(const double *part, int n, double base, double *out)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += part[i];
    out[0] = s + base;
}

__global__ void k_set_int(int *p, int v)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

// ------------------------------------------------------------------ host --
size_t store_bytes(int s) { return s == LBM_STORE_F64 ? 8 : (s == LBM_STORE_F32 ? 4 : 2); }

// bytes of one state buffer (nine planes; F16E adds its velocity plane) and
// 256 bytes of tail padding: the step's loads for the top wall node next to
// the right column read one element past the last plane (a value the wall
// routine replaces)
size_t state_bytes(int s, int N)
{
    if (s == LBM_STORE_F16E) return 2 * vel_offset(N) + 4 * (size_t)N + 256;
    if (s == LBM_STORE_F16E8) return 2 * vel_offset(N) + 2 * (size_t)N + 256;
    return store_bytes(s) * 9 * (size_t)N + 256;
}

// Q16 affine maps from the range [lo, hi] of h_q per direction: the codes
// -32767..32767 span it, h = qs i + qo.  qs is (hi - lo) / 65534 rounded up
// to 8 significant bits and qo = qs m with m the integer nearest the range's
// centre over qs, so qs m is exact in float: h = 0 (the rest state) has the
// code -m and decodes to exactly zero.  Encode: i = rint(h qi) - m, with
// qi = 1/qs and qc = 1.5 * 2^23 - m.
void q16_maps(const double *range, float *qs, float *qo, float *qi, float *qc)
{
    for (int q = 0; q < 9; ++q) {
        const double lo = range[2 * q], hi = range[2 * q + 1];
        const double ideal = (hi - lo) / 65534.0;
        int e;
        std::frexp(ideal, &e);                   // ideal = f 2^e, f in [0.5, 1)
        double sc = std::ldexp(std::ceil(std::ldexp(ideal, 8 - e)), e - 8);
        const double m = std::nearbyint(0.5 * (lo + hi) / sc);
        qs[q] = (float)sc;
        qo[q] = (float)(sc * m);
        qi[q] = (float)(1.0 / sc);
        qc[q] = (float)(12582912.0 - m);
    }
}

struct GraphKey {
    int n, parity;
    bool operator<(const GraphKey &o) const { return n != o.n ? n < o.n : parity < o.parity; }
};

}  // namespace lbm

using namespace lbm;

struct lbm_fused {
    int nx = 0, ny = 0, N = 0, storage = 0, cpt = 1, tile_cells = 256, tile_shift = 8, n_tiles = 0;
    uint32_t magic = 0;
    double om_p = 0, om_m = 0, b0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0;
    int bc[4] = {0, 0, 0, 0};
    int corners = 0;
    int n_links = 0, n_obs = 0, ring = 0, n_slots = 0, n_link_cells = 0;
    float qs[9] = {}, qo[9] = {}, qi[9] = {}, qc[9] = {};   // Q16 affine maps
    float *d_qtab = nullptr;                                // the same on the device
    void *buf[2] = {nullptr, nullptr};
    void *ckpt[2] = {nullptr, nullptr};   // the two most recent checkpoints
    int64_t ckpt_it[2] = {-1, -1};
    int *d_tile_slot = nullptr, *d_cell_off = nullptr, *d_link_cells = nullptr;
    int *d_wave_rng = nullptr, rng_stride = 0;
    uint32_t *d_run_bits = nullptr;
    uint32_t *d_cell_mask = nullptr;
    int n_perim = 0, n_perim_wg = 0;
    int blk_shift = 0;
    bool vec4 = false;
    const double *amp = nullptr;   // separable inlets (batched cases, lbm_fused_set_profiles)
    int amp_stride = 0;
    const double *prof[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    double *d_amp_own = nullptr;   // lbm_fused_set_profiles: [ring][5] amplitudes
    double *d_prof_own[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    LinkRec *d_links = nullptr;
    int *d_obs_off = nullptr;
    int8_t *d_plan_q = nullptr;
    double *d_drag_ring = nullptr, *d_drag_out = nullptr;
    int drag_out_cap = 0;
    double *d_pool = nullptr;
    int pool_cap = 0;
    int *d_inlet_off = nullptr;
    int *d_base = nullptr, *d_zero = nullptr;
    double *d_part = nullptr;
    int n_part = 0;
    hipStream_t cap_stream = nullptr;
    std::map<GraphKey, hipGraphExec_t> graphs;
    int64_t bytes = 0;
};

namespace {

template <typename P> int dmalloc(lbm_fused *h, P **p, size_t bytes)
{
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e != hipSuccess)
        return set_error(LBM_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    h->bytes += (int64_t)bytes;
    return LBM_OK;
}

Params base_params(const lbm_fused *h)
{
    Params p{};
    p.nx = h->nx;
    p.ny = h->ny;
    p.N = h->N;
    p.magic = h->magic;
    p.om_p = h->om_p;
    p.om_m = h->om_m;
    p.omf_p = (float)h->om_p;
    p.omf_m = (float)h->om_m;
    p.b0 = h->b0;
    p.a1 = h->a1;
    p.a2 = h->a2;
    p.a3 = h->a3;
    p.a4 = h->a4;
    p.bc_bottom = h->bc[0];
    p.bc_left = h->bc[1];
    p.bc_top = h->bc[2];
    p.bc_right = h->bc[3];
    p.corners = h->corners;
    p.pool = h->d_pool;
    p.inlet_off = h->d_inlet_off;
    p.amp = h->amp;
    p.amp_stride = h->amp_stride;
    for (int i = 0; i < 5; ++i) p.prof[i] = h->prof[i];
    p.ring = h->ring;
    p.base = h->d_zero;
    p.k = 0;
    p.tile_slot = h->n_slots ? h->d_tile_slot : nullptr;
    p.cell_off = h->d_cell_off;
    p.cell_mask = h->d_cell_mask;
    p.wave_rng = h->d_wave_rng;
    p.rng_stride = h->rng_stride;
    p.run_bits = h->d_run_bits;
    p.blk_shift = h->blk_shift;
    p.n_main = h->n_tiles;
    p.n_perim = h->n_perim_wg;
    p.links = h->d_links;
    p.drag_ring = h->d_drag_ring;
    p.n_links = h->n_links;
    p.link_cells = h->d_link_cells;
    p.n_link_cells = h->n_link_cells;
    p.tile_shift = h->tile_shift;
    for (int q = 0; q < 9; ++q) {
        p.qs[q] = h->qs[q];
        p.qo[q] = h->qo[q];
        p.qi[q] = h->qi[q];
        p.qc[q] = h->qc[q];
    }
    p.qtab = h->d_qtab;
    return p;
}

// Lattices whose state buffer holds 128-512 MB sweep the tiles in reverse on
// odd iterations, so that they start on the nodes the previous iteration
// wrote last, which may still sit in the 256 MB Infinity Cache (8192x2048
// obstacle array, FP16: 111.2 -> 104.8 us).  Above 512 MB of state, where
// the cache holds a small part of what was written last, one direction
// again (FP32 604 MB: 214.0 -> 212.3 us, FP64 1.2 GB: 436.4 -> 427.5).
// Smaller lattices keep one sweep direction: tile t then runs on the same XCD
// every iteration and reads from that XCD's L2 what it wrote there the
// iteration before (Turek 1073x200, FP64: 15.1 -> 13.1 us per iteration).
int sweep_rev(const lbm_fused *h, int64_t it)
{
    const size_t b = state_bytes(h->storage, h->N);
    return (b > ((size_t)128 << 20) && b <= ((size_t)512 << 20)) ? (int)(it & 1) : 0;
}

int row_of(const lbm_fused *h, int64_t it)
{
    int64_t r = it % h->ring;
    return (int)(r < 0 ? r + h->ring : r);
}

template <int S, int MODE>
void launch_fused_s(const lbm_fused *h, const Params &p, hipStream_t s)
{
    const int grid = (h->N + 511) / 512;
    k_fused<S, MODE, 2><<<grid, 256, 0, s>>>(p);
}

template <int S> void launch_step_s(const lbm_fused *h, const Params &p, hipStream_t s)
{
    const int grid = h->n_tiles + perim_groups(h->n_perim_wg);
    switch (h->cpt) {
    case 1:
        if constexpr (S == LBM_STORE_F16E8 || S == LBM_STORE_F16E) k_step_w8<S, 1><<<grid, 256, 0, s>>>(p);
        else k_step<S, 1><<<grid, 256, 0, s>>>(p);
        break;
    case 2:
        if constexpr (S == LBM_STORE_F64) k_step_w7<S, 2><<<grid, 256, 0, s>>>(p);
        else k_step<S, 2><<<grid, 256, 0, s>>>(p);
        break;
    case 4:
        if constexpr (S == LBM_STORE_F32 || S == LBM_STORE_F16 || S == LBM_STORE_BF16 ||
                      S == LBM_STORE_Q16)
            k_step_w7<S, 4><<<grid, 256, 0, s>>>(p);
        else
            k_step<S, 4><<<grid, 256, 0, s>>>(p);
        break;
    default: k_step<S, 8><<<grid, 256, 0, s>>>(p); break;
    }
}

template <int S> void launch_step_v4(const lbm_fused *h, const Params &p, hipStream_t s)
{
    const int grid = h->n_tiles + (h->n_perim_wg + 255) / 256;
    switch (h->cpt) {
    case 1: k_step_v4<S, 1><<<grid, 256, 0, s>>>(p); break;
    case 2: k_step_v4<S, 2><<<grid, 256, 0, s>>>(p); break;
    default: k_step_v4<S, 4><<<grid, 256, 0, s>>>(p); break;
    }
}

void launch_step(const lbm_fused *h, const Params &p, hipStream_t s)
{
    if (h->vec4) {
        if (h->storage == LBM_STORE_F16) launch_step_v4<LBM_STORE_F16>(h, p, s);
        else launch_step_v4<LBM_STORE_BF16>(h, p, s);
        return;
    }
    switch (h->storage) {
    case LBM_STORE_F64: launch_step_s<LBM_STORE_F64>(h, p, s); break;
    case LBM_STORE_F32: launch_step_s<LBM_STORE_F32>(h, p, s); break;
    case LBM_STORE_F16: launch_step_s<LBM_STORE_F16>(h, p, s); break;
    case LBM_STORE_BF16: launch_step_s<LBM_STORE_BF16>(h, p, s); break;
    case LBM_STORE_Q16: launch_step_s<LBM_STORE_Q16>(h, p, s); break;
    case LBM_STORE_F16E8: launch_step_s<LBM_STORE_F16E8>(h, p, s); break;
    default: launch_step_s<LBM_STORE_F16E>(h, p, s); break;
    }
}

template <int MODE> void launch_fused(const lbm_fused *h, const Params &p, hipStream_t s)
{
    switch (h->storage) {
    case LBM_STORE_F64: launch_fused_s<LBM_STORE_F64, MODE>(h, p, s); break;
    case LBM_STORE_F32: launch_fused_s<LBM_STORE_F32, MODE>(h, p, s); break;
    case LBM_STORE_F16: launch_fused_s<LBM_STORE_F16, MODE>(h, p, s); break;
    case LBM_STORE_BF16: launch_fused_s<LBM_STORE_BF16, MODE>(h, p, s); break;
    case LBM_STORE_Q16: launch_fused_s<LBM_STORE_Q16, MODE>(h, p, s); break;
    case LBM_STORE_F16E8: launch_fused_s<LBM_STORE_F16E8, MODE>(h, p, s); break;
    default: launch_fused_s<LBM_STORE_F16E, MODE>(h, p, s); break;
    }
}

void launch_drag_only(const lbm_fused *h, const Params &p, hipStream_t s)
{
    const int grid = (h->n_link_cells + 255) / 256;
    if (grid == 0) return;
    switch (h->storage) {
    case LBM_STORE_F64: k_drag_only<LBM_STORE_F64><<<grid, 256, 0, s>>>(p); break;
    case LBM_STORE_F32: k_drag_only<LBM_STORE_F32><<<grid, 256, 0, s>>>(p); break;
    case LBM_STORE_F16: k_drag_only<LBM_STORE_F16><<<grid, 256, 0, s>>>(p); break;
    case LBM_STORE_BF16: k_drag_only<LBM_STORE_BF16><<<grid, 256, 0, s>>>(p); break;
    case LBM_STORE_Q16: k_drag_only<LBM_STORE_Q16><<<grid, 256, 0, s>>>(p); break;
    case LBM_STORE_F16E8: k_drag_only<LBM_STORE_F16E8><<<grid, 256, 0, s>>>(p); break;
    default: k_drag_only<LBM_STORE_F16E><<<grid, 256, 0, s>>>(p); break;
    }
}

inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Builds the node-sorted link records and the per-tile offset tables.
int build_links(lbm_fused *h, const lbm_fused_desc *d)
{
    const int nx = h->nx, ny = h->ny, nl = d->n_links;
    std::vector<LinkRec> recs(nl);
    std::vector<int8_t> plan_q(nl);
    auto wrap = [](long long k, int n, long long &o) {
        if (k < -(long long)n || k >= n) return false;
        o = k < 0 ? k + n : k;
        return true;
    };
    for (int k = 0; k < nl; ++k) {
        const long long ri = d->links_host[3 * k], rj = d->links_host[3 * k + 1];
        const long long q = d->links_host[3 * k + 2];
        long long i, j;
        if (!wrap(ri, nx, i) || !wrap(rj, ny, j) || q < 1 || q > 8)
            return set_error(LBM_ERR_ARG, "link " + std::to_string(k) + " out of range");
        LinkRec L{};
        L.cell = (int)(i * ny + j);
        L.q = (int8_t)q;
        L.qb = (int8_t)NS[q];
        L.dslot = k;
        L.mode = d->link_ibb_host ? (d->link_ibb_host[k] ? 1 : 0) : 0;
        plan_q[k] = (int8_t)q;
        if (L.mode) {
            const int qb = NS[q];
            long long im, jm, imm, jmm;
            // cb = c[qb] (nb.py:89-93); indices wrap like NumPy's
            if (!wrap(ri + CX[qb], nx, im) || !wrap(rj + CY[qb], ny, jm) ||
                !wrap(ri + 2 * CX[qb], nx, imm) || !wrap(rj + 2 * CY[qb], ny, jmm))
                return set_error(LBM_ERR_ARG,
                                 "IBB stencil of link " + std::to_string(k) + " leaves the lattice");
            const double p = d->ibb_host[k];
            const double pp = 2.0 * p;
            if (p < 0.5) {  // nb.py:97-100
                L.planeB = (int8_t)q;
                L.cellB = (int)(im * ny + jm);
                L.planeC = (int8_t)q;
                L.cellC = (int)(imm * ny + jmm);
                L.c0 = p * (pp + 1.0);
                L.c1 = (1.0 + pp) * (1.0 - pp);
                L.c2 = -(p * (1.0 - pp));
            } else {        // nb.py:101-104
                L.planeB = (int8_t)qb;
                L.cellB = L.cell;
                L.planeC = (int8_t)qb;
                L.cellC = (int)(im * ny + jm);
                L.c0 = 1.0 / (p * (pp + 1.0));
                L.c1 = (pp - 1.0) / p;
                L.c2 = (1.0 - pp) / (1.0 + pp);
            }
        }
        recs[k] = L;
    }
    if (!d->obstacles_first) {
        for (const auto &L : recs) {
            const int x = L.cell / ny, y = L.cell % ny;
            if (x == 0 || x == nx - 1 || y == 0 || y == ny - 1)
                return set_error(LBM_ERR_UNSUPPORTED,
                                 "obstacle links on the perimeter applied after the walls");
        }
    }
    std::stable_sort(recs.begin(), recs.end(),
                     [](const LinkRec &a, const LinkRec &b) { return a.cell < b.cell; });
    std::vector<int> tile_slot(h->n_tiles, -1);
    std::vector<int> cell_off;
    std::vector<int> link_cells;
    const int tc = h->tile_cells;
    int slot = 0;
    for (int k = 0; k < nl;) {
        const int tile = recs[k].cell >> h->tile_shift;
        tile_slot[tile] = slot;
        const size_t base = cell_off.size();
        cell_off.resize(base + tc + 1, 0);
        int k2 = k;
        std::vector<int> cnt(tc, 0);
        while (k2 < nl && (recs[k2].cell >> h->tile_shift) == tile) {
            ++cnt[recs[k2].cell & (tc - 1)];
            if (link_cells.empty() || link_cells.back() != recs[k2].cell)
                link_cells.push_back(recs[k2].cell);
            ++k2;
        }
        int acc = k;
        for (int c = 0; c < tc; ++c) {
            cell_off[base + c] = acc;
            acc += cnt[c];
        }
        cell_off[base + tc] = acc;
        k = k2;
        ++slot;
    }
    h->n_slots = slot;
    h->n_link_cells = (int)link_cells.size();
    // one bit per node of each link tile: the node has links
    std::vector<uint32_t> mask((size_t)std::max(slot, 1) * (tc / 32), 0u);
    for (const auto &L : recs) {
        const int sl = tile_slot[L.cell >> h->tile_shift];
        const int local = L.cell & (tc - 1);
        mask[(size_t)sl * (tc / 32) + (local >> 5)] |= 1u << (local & 31);
    }
    int e;
    if ((e = dmalloc(h, &h->d_tile_slot, sizeof(int) * h->n_tiles))) return e;
    if ((e = dmalloc(h, &h->d_cell_off, sizeof(int) * cell_off.size()))) return e;
    if ((e = dmalloc(h, &h->d_links, sizeof(LinkRec) * nl))) return e;
    if ((e = dmalloc(h, &h->d_link_cells, sizeof(int) * link_cells.size()))) return e;
    if ((e = dmalloc(h, &h->d_plan_q, nl))) return e;
    if ((e = dmalloc(h, &h->d_cell_mask, sizeof(uint32_t) * mask.size()))) return e;
    LBM_HIP_TRY(hipMemcpy(h->d_cell_mask, mask.data(), sizeof(uint32_t) * mask.size(),
                          hipMemcpyHostToDevice));
    LBM_HIP_TRY(hipMemcpy(h->d_tile_slot, tile_slot.data(), sizeof(int) * h->n_tiles,
                          hipMemcpyHostToDevice));
    if (!cell_off.empty())
        LBM_HIP_TRY(hipMemcpy(h->d_cell_off, cell_off.data(), sizeof(int) * cell_off.size(),
                              hipMemcpyHostToDevice));
    if (slot > 0 && h->storage != LBM_STORE_F64) {
        // the record bounds of every 64-node run, per tile, padded to a power
        // of two: one scalar load pair per wave, no dependent slot lookup
        const int runs = tc / 64;
        int rs = 1;
        while (rs < runs + 1) rs <<= 1;
        h->rng_stride = rs;
        std::vector<int> rng((size_t)h->n_tiles * rs, 0);
        for (int t = 0; t < h->n_tiles; ++t) {
            if (tile_slot[t] < 0) continue;
            for (int j = 0; j <= runs; ++j)
                rng[(size_t)t * rs + j] = cell_off[(size_t)tile_slot[t] * (tc + 1) + j * 64];
        }
        // one bit per run: the table a wave reads first (4 KiB per 2^20 nodes,
        // held in the L2), so that only the runs with links read their row
        std::vector<uint32_t> bits(((size_t)h->n_tiles * runs + 31) / 32, 0u);
        for (int t = 0; t < h->n_tiles; ++t)
            for (int j = 0; j < runs; ++j) {
                const size_t r = (size_t)t * runs + j;
                if (rng[(size_t)t * rs + j + 1] > rng[(size_t)t * rs + j])
                    bits[r >> 5] |= 1u << (r & 31);
            }
        if ((e = dmalloc(h, &h->d_wave_rng, sizeof(int) * rng.size()))) return e;
        if ((e = dmalloc(h, &h->d_run_bits, sizeof(uint32_t) * bits.size()))) return e;
        LBM_HIP_TRY(hipMemcpy(h->d_wave_rng, rng.data(), sizeof(int) * rng.size(),
                              hipMemcpyHostToDevice));
        LBM_HIP_TRY(hipMemcpy(h->d_run_bits, bits.data(), sizeof(uint32_t) * bits.size(),
                              hipMemcpyHostToDevice));
    }
    if (nl) {
        LBM_HIP_TRY(hipMemcpy(h->d_links, recs.data(), sizeof(LinkRec) * nl, hipMemcpyHostToDevice));
        LBM_HIP_TRY(hipMemcpy(h->d_plan_q, plan_q.data(), nl, hipMemcpyHostToDevice));
        LBM_HIP_TRY(hipMemcpy(h->d_link_cells, link_cells.data(), sizeof(int) * link_cells.size(),
                              hipMemcpyHostToDevice));
    }
    // obstacle link ranges
    std::vector<int> obs_off(d->n_obs + 1, 0);
    if (d->n_obs > 0) {
        for (int o = 0; o <= d->n_obs; ++o) obs_off[o] = d->obs_offsets_host[o];
        if (obs_off[0] != 0 || obs_off[d->n_obs] != nl)
            return set_error(LBM_ERR_ARG, "obs_offsets must span [0, n_links]");
        for (int o = 0; o < d->n_obs; ++o)
            if (obs_off[o + 1] < obs_off[o]) return set_error(LBM_ERR_ARG, "obs_offsets decrease");
    }
    if ((e = dmalloc(h, &h->d_obs_off, sizeof(int) * obs_off.size()))) return e;
    LBM_HIP_TRY(hipMemcpy(h->d_obs_off, obs_off.data(), sizeof(int) * obs_off.size(),
                          hipMemcpyHostToDevice));
    return LBM_OK;
}

void free_all(lbm_fused *h)
{
    for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
    h->graphs.clear();
    void *ptrs[] = {h->buf[0], h->buf[1], h->ckpt[0], h->ckpt[1], h->d_tile_slot, h->d_cell_off, h->d_cell_mask,
                    h->d_wave_rng, h->d_run_bits,
                    h->d_link_cells, h->d_links, h->d_obs_off, h->d_plan_q, h->d_drag_ring,
                    h->d_drag_out, h->d_pool, h->d_inlet_off, h->d_base, h->d_zero, h->d_part,
                    h->d_qtab, h->d_amp_own, h->d_prof_own[0], h->d_prof_own[1],
                    h->d_prof_own[2], h->d_prof_own[3], h->d_prof_own[4]};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (h->cap_stream) (void)hipStreamDestroy(h->cap_stream);
}

}  // namespace

extern "C" int lbm_abi_version(void) { return LBM_ABI_VERSION; }

extern "C" const char *lbm_last_error(void) { return lbm::g_last_error.c_str(); }

extern "C" int64_t lbm_fused_device_bytes(const lbm_fused *h) { return h ? h->bytes : 0; }

// ------------------------------------------------ allocation-free drag read --
// Staging array of lbm_drag_read_noalloc: the per-obstacle sums go here
// instead of the buffer lbm_drag_read allocates and grows.  One copy per GPU,
// held until the process ends.  Host code never passes `scratch` itself (on
// the host the name is a stand-in, not a device address): it passes the
// address hipGetSymbolAddress returns for the current device, kept per device.
static __device__ __align__(256) double scratch[LBM_NOALLOC_DRAG_MAX];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};

// Address of `scratch` on the current device, looked up once per device
// (lbm_fused_create, or the first lbm_drag_read_noalloc call on a device whose
// handles all came from a batch).
int scratch_address(double **out)
{
    int dev = 0;
    LBM_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_error(LBM_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_error(LBM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                              hipGetErrorString(e));
        g_scratch_dev[dev] = static_cast<double *>(ptr);
    }
    *out = g_scratch_dev[dev];
    return LBM_OK;
}
}  // namespace

extern "C" int lbm_fused_create(const lbm_fused_desc *d, lbm_fused **out)
{
    if (!d || !out) return set_error(LBM_ERR_ARG, "lbm_fused_create: null argument");
    *out = nullptr;
    if (d->nx < 4 || d->ny < 4) return set_error(LBM_ERR_ARG, "nx and ny must be >= 4");
    if ((long long)d->nx * d->ny * 9 >= (1LL << 31))
        return set_error(LBM_ERR_ARG, "lattice too large for 32-bit population indices");
    if (d->storage < 0 || d->storage > LBM_STORE_F16E8)
        return set_error(LBM_ERR_ARG, "bad storage");
    if (d->storage == LBM_STORE_Q16) {
        if (!d->q16_range_host)
            return set_error(LBM_ERR_ARG, "Q16 storage needs q16_range_host");
        for (int q = 0; q < 9; ++q) {
            const double lo = d->q16_range_host[2 * q], hi = d->q16_range_host[2 * q + 1];
            if (!(hi > lo) || !(hi - lo < 1.0) || !(lo > -1.0) || !(hi < 1.0))
                return set_error(LBM_ERR_ARG, "Q16 range of direction " + std::to_string(q) +
                                                  " must satisfy -1 < lo < hi < 1");
        }
    }
    if (d->ring < 2) return set_error(LBM_ERR_ARG, "ring must be >= 2");
    if (d->n_links < 0 || (d->n_links > 0 && (!d->links_host || !d->ibb_host)))
        return set_error(LBM_ERR_ARG, "bad link arrays");
    if (d->n_obs < 0 || (d->n_obs > 0 && !d->obs_offsets_host))
        return set_error(LBM_ERR_ARG, "bad obstacle ranges");
    // The fused plan: the four Zou-He walls (right wall velocity or pressure)
    // followed by the four corners, as every reference app's set_bc does.
    if (d->bc[0] != LBM_BC_VELOCITY || d->bc[1] != LBM_BC_VELOCITY ||
        d->bc[2] != LBM_BC_VELOCITY ||
        (d->bc[3] != LBM_BC_VELOCITY && d->bc[3] != LBM_BC_PRESSURE) || !d->corners)
        return set_error(LBM_ERR_UNSUPPORTED,
                         "fused plan needs Zou-He velocity walls on bottom/left/top, velocity "
                         "or pressure on the right, and the four corners");
    auto h = std::make_unique<lbm_fused>();
    h->nx = d->nx;
    h->ny = d->ny;
    h->N = d->nx * d->ny;
    h->storage = d->storage;
    h->cpt = d->cells_per_thread;
    if (d->storage == LBM_STORE_Q16) q16_maps(d->q16_range_host, h->qs, h->qo, h->qi, h->qc);
    // default nodes per work-item, the fastest on the 8192x2048 obstacle array:
    // FP16, BF16 and Q16 2 (FP16 109.4 us against 114.4 at 4 and 120.5 at 1,
    // Q16 116.7 against 133.7 at 4: at 2 nodes the 16-bit tiles fit 64 VGPRs,
    // 8 waves per SIMD), FP64, FP32, f16e and f16e8 1 (FP64 433 against 452
    // at 2, FP32 220 against 226.5, f16e8 149.6 against 149.5)
    if (h->cpt == 0)
        h->cpt = (d->storage == LBM_STORE_F16 || d->storage == LBM_STORE_BF16 ||
                  d->storage == LBM_STORE_Q16) ? 2 : 1;
    if (h->cpt != 1 && h->cpt != 2 && h->cpt != 4 && h->cpt != 8)
        return set_error(LBM_ERR_ARG, "cells_per_thread must be 1, 2, 4 or 8");
    // 16-bit forms: four nodes per work-item on request, when the lattice
    // allows it (one node per work-item with four per tile row is faster on
    // the 8192x2048 obstacle array: 122 against 111 GLUPS)
    const bool sixteen = d->storage == LBM_STORE_F16 || d->storage == LBM_STORE_BF16;
    if (sixteen && d->ny % 4 == 0 && d->vector_width == 4) {
        const int cpt = (d->cells_per_thread == 0 || h->cpt >= 4) ? 4 : h->cpt;
        if ((long long)d->nx * d->ny % (1024LL * cpt) == 0) {
            h->vec4 = true;
            h->cpt = cpt;
        }
    }
    const int per_thread = h->vec4 ? 4 * h->cpt : h->cpt;
    h->tile_cells = 256 * per_thread;
    h->tile_shift = 8 + (per_thread == 1 ? 0 : per_thread == 2 ? 1 : per_thread == 4 ? 2
                         : per_thread == 8 ? 3 : 4);
    h->n_tiles = (h->N + h->tile_cells - 1) / h->tile_cells;
    // Tiles dealt to the XCDs in blocks of consecutive tiles.  Lattices of
    // up to 128 MB of state (one sweep direction, each tile on the same XCD
    // every iteration): blocks of 4, so that a tile's neighbours share its
    // XCD's L2 (Turek 1073x200 FP64 through the drop-in run: 14.3 -> 18.2
    // GLUPS).  Larger ones, float forms: the tiles of one column on one XCD,
    // the columns round-robin over the XCDs (a power-of-two number of tiles
    // per column; on the 8192x2048 obstacle array FP16 106.6 -> 105.5 us,
    // f16e8 142.5 -> 139.2, f16e 157.2 -> 152.0; FP64 430.8 -> 450.5, so
    // FP64 keeps the dispatch order there).
    if (state_bytes(d->storage, h->N) <= ((size_t)128 << 20)) {
        h->blk_shift = 2;
    } else if (d->storage != LBM_STORE_F64 && d->ny % h->tile_cells == 0) {
        const int tpc = d->ny / h->tile_cells;
        if ((tpc & (tpc - 1)) == 0)
            while ((1 << h->blk_shift) < tpc) ++h->blk_shift;
    }
    h->n_perim = 2 * d->nx + 2 * d->ny - 4;
    // nodes of the perimeter work-groups: the left and right columns (corners
    // included) where the one-node tiles complete their bottom and top wall
    // nodes (16-bit forms and f16e: ny >= 64), every
    // perimeter node otherwise
    const bool tile_walls = !h->vec4 && (d->storage == LBM_STORE_F64 ||
                                         d->storage == LBM_STORE_F32 ||
                                         (is16(d->storage) && d->ny >= 64));
    h->n_perim_wg = tile_walls ? 2 * d->ny : h->n_perim;
    h->magic = (uint32_t)((1ULL << 32) / (uint64_t)d->ny) + 1u;
    h->om_p = d->om_p;
    h->om_m = d->om_m;
    // the scalar factors of nb.py:25 and nb.py:32-35
    h->b0 = 1.0 - d->om_p;
    h->a1 = 1.0 - 0.5 * (d->om_p + d->om_m);
    h->a2 = 0.5 * (d->om_p - d->om_m);
    h->a3 = 0.5 * (d->om_p + d->om_m);
    h->a4 = 0.5 * (d->om_p - d->om_m);
    for (int i = 0; i < 4; ++i) h->bc[i] = d->bc[i];
    h->corners = d->corners;
    h->n_links = d->n_links;
    h->n_obs = d->n_obs;
    h->ring = d->ring;
    int e;
    const size_t sb = state_bytes(d->storage, h->N);
    if ((e = dmalloc(h.get(), &h->buf[0], sb)) || (e = dmalloc(h.get(), &h->buf[1], sb))) {
        free_all(h.get());
        return e;
    }
    if ((e = build_links(h.get(), d))) {
        free_all(h.get());
        return e;
    }
    if ((e = dmalloc(h.get(), &h->d_drag_ring, sizeof(double) * (size_t)h->ring *
                                                   (h->n_links ? h->n_links : 1))) ||
        (e = dmalloc(h.get(), &h->d_inlet_off, sizeof(int) * 5 * (size_t)h->ring)) ||
        (e = dmalloc(h.get(), &h->d_base, sizeof(int))) ||
        (e = dmalloc(h.get(), &h->d_zero, sizeof(int)))) {
        free_all(h.get());
        return e;
    }
    if (hipMemset(h->d_inlet_off, 0, sizeof(int) * 5 * (size_t)h->ring) != hipSuccess ||
        hipMemset(h->d_zero, 0, sizeof(int)) != hipSuccess ||
        hipMemset(h->d_base, 0, sizeof(int)) != hipSuccess) {
        free_all(h.get());
        return set_error(LBM_ERR_HIP, "hipMemset failed");
    }
    if (d->storage == LBM_STORE_Q16) {
        float tab[36];
        for (int q = 0; q < 9; ++q) {
            tab[q] = h->qs[q];
            tab[9 + q] = h->qo[q];
            tab[18 + q] = h->qi[q];
            tab[27 + q] = h->qc[q];
        }
        if ((e = dmalloc(h.get(), &h->d_qtab, sizeof(tab)))) {
            free_all(h.get());
            return e;
        }
        if (hipMemcpy(h->d_qtab, tab, sizeof(tab), hipMemcpyHostToDevice) != hipSuccess) {
            free_all(h.get());
            return set_error(LBM_ERR_HIP, "hipMemcpy failed");
        }
    }
    double *stage = nullptr;   // the drag staging array's address on this device
    if ((e = scratch_address(&stage))) {
        free_all(h.get());
        return e;
    }
    *out = h.release();
    return LBM_OK;
}

extern "C" void lbm_fused_destroy(lbm_fused *h)
{
    if (!h) return;
    (void)hipDeviceSynchronize();
    free_all(h);
    delete h;
}

extern "C" int lbm_fused_set_inlets(lbm_fused *h, int row0, int nrows, const int32_t *off_host,
                                    int pool_len, const double *pool_host, void *stream)
{
    if (!h || nrows < 0 || nrows > h->ring || pool_len < 0 || (nrows && !off_host) ||
        (pool_len && !pool_host))
        return set_error(LBM_ERR_ARG, "lbm_fused_set_inlets: bad argument");
    // every side array of every row lies inside the pool: u_bot and u_top
    // [2][nx], u_left and u_right [2][ny], rho_right [ny]
    const int side_len[5] = {2 * h->nx, 2 * h->ny, 2 * h->nx, 2 * h->ny, h->ny};
    for (int i = 0; i < 5 * nrows; ++i) {
        const int64_t o = off_host[i];
        if (o < 0 || o + side_len[i % 5] > (int64_t)pool_len)
            return set_error(LBM_ERR_ARG, "inlet row " + std::to_string(i / 5) + ", side " +
                                              std::to_string(i % 5) + ": offset " +
                                              std::to_string(o) + " + length " +
                                              std::to_string(side_len[i % 5]) +
                                              " exceeds the pool (" + std::to_string(pool_len) +
                                              " doubles)");
    }
    // the previous contents may still be read by queued launches
    LBM_HIP_TRY(hipStreamSynchronize(S(stream)));
    // rows never written keep offset 0: the allocation holds at least one
    // side array of the longest kind there (zeros until a pool covers it)
    const int cap_need = std::max(pool_len, 2 * std::max(h->nx, h->ny));
    if (cap_need > h->pool_cap) {
        if (h->d_pool) {
            (void)hipFree(h->d_pool);
            h->bytes -= (int64_t)sizeof(double) * h->pool_cap;
        }
        h->d_pool = nullptr;
        h->pool_cap = 0;
        if (int e = dmalloc(h, &h->d_pool, sizeof(double) * cap_need)) return e;
        h->pool_cap = cap_need;
        LBM_HIP_TRY(hipMemset(h->d_pool, 0, sizeof(double) * cap_need));
        for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
        h->graphs.clear();
    }
    if (pool_len)
        LBM_HIP_TRY(hipMemcpy(h->d_pool, pool_host, sizeof(double) * pool_len,
                              hipMemcpyHostToDevice));
    // rows row0 .. row0+nrows-1 of the ring: at most two contiguous pieces
    int r = 0;
    while (r < nrows) {
        const int row = row_of(h, (int64_t)row0 + r);
        const int len = std::min(nrows - r, h->ring - row);
        LBM_HIP_TRY(hipMemcpy(h->d_inlet_off + 5 * row, off_host + 5 * r, 5 * sizeof(int) * len,
                              hipMemcpyHostToDevice));
        r += len;
    }
    return LBM_OK;
}

extern "C" int lbm_fused_set_profiles(lbm_fused *h, const double *const *profile_host,
                                      void *stream)
{
    if (!h || !profile_host) return set_error(LBM_ERR_ARG, "lbm_fused_set_profiles: bad argument");
    if (h->amp && !h->d_amp_own)
        return set_error(LBM_ERR_ARG, "lbm_fused_set_profiles: the wall inputs of a batch case "
                                      "are the batch's (lbm_batch_set_amplitudes)");
    // queued launches may still read the previous profiles
    LBM_HIP_TRY(hipStreamSynchronize(S(stream)));
    const int len[5] = {2 * h->nx, 2 * h->ny, 2 * h->nx, 2 * h->ny, h->ny};
    if (!h->d_amp_own) {
        if (int e = dmalloc(h, &h->d_amp_own, sizeof(double) * 5 * (size_t)h->ring)) return e;
        LBM_HIP_TRY(hipMemset(h->d_amp_own, 0, sizeof(double) * 5 * (size_t)h->ring));
        for (int sd = 0; sd < 5; ++sd)
            if (int e = dmalloc(h, &h->d_prof_own[sd], sizeof(double) * len[sd])) return e;
    }
    for (int sd = 0; sd < 5; ++sd) {
        if (profile_host[sd])
            LBM_HIP_TRY(hipMemcpy(h->d_prof_own[sd], profile_host[sd], sizeof(double) * len[sd],
                                  hipMemcpyHostToDevice));
        else
            LBM_HIP_TRY(hipMemset(h->d_prof_own[sd], 0, sizeof(double) * len[sd]));
        h->prof[sd] = h->d_prof_own[sd];
    }
    h->amp = h->d_amp_own;
    h->amp_stride = 5;
    // captured launches hold the previous inlet pointers
    for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
    h->graphs.clear();
    return LBM_OK;
}

extern "C" int lbm_fused_set_amplitudes(lbm_fused *h, int64_t row0, int nrows,
                                        const double *amp_host, void *stream)
{
    if (!h || row0 < 0 || nrows < 0 || nrows > h->ring || (nrows && !amp_host))
        return set_error(LBM_ERR_ARG, "lbm_fused_set_amplitudes: bad argument");
    if (!h->d_amp_own)
        return set_error(LBM_ERR_ARG, "lbm_fused_set_amplitudes: call lbm_fused_set_profiles first");
    // queued launches may still read the rows being replaced
    LBM_HIP_TRY(hipStreamSynchronize(S(stream)));
    int r = 0;
    while (r < nrows) {
        const int row = row_of(h, row0 + r);
        const int len = std::min(nrows - r, h->ring - row);
        LBM_HIP_TRY(hipMemcpy(h->d_amp_own + 5 * (size_t)row, amp_host + 5 * (size_t)r,
                              5 * sizeof(double) * len, hipMemcpyHostToDevice));
        r += len;
    }
    return LBM_OK;
}

extern "C" int lbm_init(lbm_fused *h, const double *g, double *rho, double *u, void *stream)
{
    if (!h || ((rho == nullptr) != (u == nullptr)))
        return set_error(LBM_ERR_ARG, "lbm_init: bad argument");
    Params p = base_params(h);
    p.g0 = g;
    p.dst = h->buf[0];
    p.rho_out = rho;
    p.u_out = u;
    launch_fused<MODE_INIT>(h, p, S(stream));
    LBM_LAUNCH_CHECK();
    return LBM_OK;
}

extern "C" int lbm_step(lbm_fused *h, int64_t it, double *rho, double *u, void *stream)
{
    if (!h || it < 1 || ((rho == nullptr) != (u == nullptr)))
        return set_error(LBM_ERR_ARG, "lbm_step: bad argument");
    if (!h->d_pool && !h->amp) return set_error(LBM_ERR_ARG, "lbm_step: no inlet table");
    Params p = base_params(h);
    p.src = h->buf[(it - 1) & 1];
    p.dst = h->buf[it & 1];
    p.k = row_of(h, it - 1);
    p.rho_out = rho;
    p.u_out = u;
    p.rev = sweep_rev(h, it);
    launch_step(h, p, S(stream));
    LBM_LAUNCH_CHECK();
    return LBM_OK;
}

namespace {
// The graph of n steps starting at an iteration of it0's parity, captured
// and instantiated on first use.
int step_graph(lbm_fused *h, int64_t it0, int n, hipGraphExec_t *out)
{
    const GraphKey key{n, (int)(it0 & 1)};
    auto itg = h->graphs.find(key);
    if (itg == h->graphs.end()) {
        if (!h->cap_stream)   // created on first use: batched cases never capture
            LBM_HIP_TRY(hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking));
        hipGraph_t graph;
        LBM_HIP_TRY(hipStreamBeginCapture(h->cap_stream, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < n; ++k) {
            Params p = base_params(h);
            p.src = h->buf[(it0 + k - 1) & 1];
            p.dst = h->buf[(it0 + k) & 1];
            p.base = h->d_base;
            p.k = k;
            p.rev = sweep_rev(h, it0 + k);
            launch_step(h, p, h->cap_stream);
        }
        hipError_t ec = hipStreamEndCapture(h->cap_stream, &graph);
        if (ec != hipSuccess)
            return set_error(LBM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
        hipGraphExec_t exec;
        ec = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ec != hipSuccess)
            return set_error(LBM_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ec));
        itg = h->graphs.emplace(key, exec).first;
    }
    *out = itg->second;
    return LBM_OK;
}
}  // namespace

extern "C" int lbm_steps(lbm_fused *h, int64_t it0, int n, void *stream)
{
    if (!h || it0 < 1 || n < 0) return set_error(LBM_ERR_ARG, "lbm_steps: bad argument");
    if (n == 0) return LBM_OK;
    if (!h->d_pool && !h->amp) return set_error(LBM_ERR_ARG, "lbm_steps: no inlet table");
    hipGraphExec_t exec;
    if (int e = step_graph(h, it0, n, &exec)) return e;
    k_set_int<<<1, 64, 0, S(stream)>>>(h->d_base, row_of(h, it0 - 1));
    LBM_LAUNCH_CHECK();
    LBM_HIP_TRY(hipGraphLaunch(exec, S(stream)));
    return LBM_OK;
}

extern "C" int lbm_steps_prepare(lbm_fused *h, int64_t it0, int n, void *stream)
{
    if (!h || it0 < 1 || n < 0) return set_error(LBM_ERR_ARG, "lbm_steps_prepare: bad argument");
    if (n == 0) return LBM_OK;
    if (!h->d_pool && !h->amp) return set_error(LBM_ERR_ARG, "lbm_steps_prepare: no inlet table");
    hipGraphExec_t exec;
    if (int e = step_graph(h, it0, n, &exec)) return e;
    LBM_HIP_TRY(hipGraphUpload(exec, S(stream)));
    return LBM_OK;
}

extern "C" int lbm_epilogue(lbm_fused *h, int64_t it, double *g, double *gup, double *rho,
                            double *u, void *stream)
{
    if (!h || it < 0 || !g || !gup || ((rho == nullptr) != (u == nullptr)))
        return set_error(LBM_ERR_ARG, "lbm_epilogue: bad argument");
    if (!h->d_pool && !h->amp) return set_error(LBM_ERR_ARG, "lbm_epilogue: no inlet table");
    Params p = base_params(h);
    p.src = h->buf[it & 1];
    p.g_out = g;
    p.gup_out = gup;
    p.k = row_of(h, it);
    p.rho_out = rho;
    p.u_out = u;
    launch_fused<MODE_EPILOGUE>(h, p, S(stream));
    LBM_LAUNCH_CHECK();
    return LBM_OK;
}

extern "C" int lbm_drag_only(lbm_fused *h, int64_t it, void *stream)
{
    if (!h || it < 0) return set_error(LBM_ERR_ARG, "lbm_drag_only: bad argument");
    if (!h->d_pool && !h->amp) return set_error(LBM_ERR_ARG, "lbm_drag_only: no inlet table");
    Params p = base_params(h);
    p.src = h->buf[it & 1];
    p.k = row_of(h, it);
    launch_drag_only(h, p, S(stream));
    LBM_LAUNCH_CHECK();
    return LBM_OK;
}

extern "C" int lbm_drag_read(lbm_fused *h, int64_t it0, int n, double *out, void *stream)
{
    if (!h || n < 0 || n > h->ring || it0 < 0 || (n && !out))
        return set_error(LBM_ERR_ARG, "lbm_drag_read: bad argument");
    if (n == 0 || h->n_obs == 0) return LBM_OK;
    const int need = n * h->n_obs * 2;
    if (need > h->drag_out_cap) {
        LBM_HIP_TRY(hipStreamSynchronize(S(stream)));
        if (h->d_drag_out) {
            (void)hipFree(h->d_drag_out);
            h->bytes -= (int64_t)sizeof(double) * h->drag_out_cap;
        }
        h->d_drag_out = nullptr;
        h->drag_out_cap = 0;
        if (int e = dmalloc(h, &h->d_drag_out, sizeof(double) * need)) return e;
        h->drag_out_cap = need;
    }
    const int threads = n * h->n_obs;
    k_drag_reduce<<<(threads + 127) / 128, 128, 0, S(stream)>>>(
        h->d_drag_ring, h->n_links, h->ring, it0, n, h->d_obs_off, h->n_obs, h->d_plan_q,
        h->d_drag_out);
    LBM_LAUNCH_CHECK();
    LBM_HIP_TRY(hipMemcpyAsync(out, h->d_drag_out, sizeof(double) * need, hipMemcpyDeviceToHost,
                               S(stream)));
    LBM_HIP_TRY(hipStreamSynchronize(S(stream)));
    return LBM_OK;
}

extern "C" int lbm_drag_read_noalloc(lbm_fused *h, int64_t it0, int n, double *out, void *stream)
{
    if (!h || n < 0 || n > h->ring || it0 < 0 || (n && !out))
        return set_error(LBM_ERR_ARG, "lbm_drag_read_noalloc: bad argument");
    if (n == 0 || h->n_obs == 0) return LBM_OK;
    const int64_t need = (int64_t)n * h->n_obs * 2;
    if (need > LBM_NOALLOC_DRAG_MAX)
        return set_error(LBM_ERR_ARG,
                         "lbm_drag_read_noalloc: n * n_obs * 2 = " + std::to_string(need) +
                             " exceeds the staging array (LBM_NOALLOC_DRAG_MAX = " +
                             std::to_string(LBM_NOALLOC_DRAG_MAX) +
                             " doubles); read fewer iterations per call");
    double *stage = nullptr;
    if (int e = scratch_address(&stage)) return e;
    // k_drag_reduce writes every one of the `need` sums: nothing to clear
    const int threads = n * h->n_obs;
    k_drag_reduce<<<(threads + 127) / 128, 128, 0, S(stream)>>>(
        h->d_drag_ring, h->n_links, h->ring, it0, n, h->d_obs_off, h->n_obs, h->d_plan_q, stage);
    LBM_LAUNCH_CHECK();
    LBM_HIP_TRY(hipMemcpyAsync(out, stage, sizeof(double) * need, hipMemcpyDeviceToHost,
                               S(stream)));
    LBM_HIP_TRY(hipStreamSynchronize(S(stream)));
    return LBM_OK;
}

extern "C" int lbm_fused_save(lbm_fused *h, int64_t it, void *stream)
{
    if (!h || it < 0) return set_error(LBM_ERR_ARG, "lbm_fused_save: bad argument");
    const size_t sb = state_bytes(h->storage, h->N);
    // the slot already holding iteration it, else the older of the two
    int k = h->ckpt_it[0] == it ? 0 : h->ckpt_it[1] == it ? 1
                                      : (h->ckpt_it[0] <= h->ckpt_it[1] ? 0 : 1);
    if (!h->ckpt[k])
        if (int e = dmalloc(h, &h->ckpt[k], sb)) return e;
    LBM_HIP_TRY(hipMemcpyAsync(h->ckpt[k], h->buf[it & 1], sb, hipMemcpyDeviceToDevice,
                               S(stream)));
    h->ckpt_it[k] = it;
    return LBM_OK;
}

extern "C" int lbm_fused_restore(lbm_fused *h, int64_t it, void *stream)
{
    const int k = !h ? -1 : h->ckpt_it[0] == it ? 0 : h->ckpt_it[1] == it ? 1 : -1;
    if (!h || it < 0 || k < 0 || !h->ckpt[k])
        return set_error(LBM_ERR_ARG, "lbm_fused_restore: no checkpoint of that iteration");
    const size_t sb = state_bytes(h->storage, h->N);
    LBM_HIP_TRY(hipMemcpyAsync(h->buf[it & 1], h->ckpt[k], sb, hipMemcpyDeviceToDevice,
                               S(stream)));
    return LBM_OK;
}

namespace {
// N * sum_q w_q for the float forms (their partial sums hold the deviations)
double mass_base(int storage, int N)
{
    if (storage == LBM_STORE_F64) return 0.0;
    double ws = 0.0;
    for (int q = 0; q < 9; ++q) ws += weight<double>(q);
    return (double)N * ws;
}
}  // namespace

extern "C" int lbm_mass(lbm_fused *h, int64_t it, double *out, void *stream)
{
    if (!h || it < 0 || !out) return set_error(LBM_ERR_ARG, "lbm_mass: bad argument");
    const int nb = 1024;
    if (!h->d_part) {
        if (int e = dmalloc(h, &h->d_part, sizeof(double) * (nb + 1))) return e;
        h->n_part = nb;
    }
    Params p = base_params(h);
    p.src = h->buf[it & 1];
    double *part = h->d_part;
    switch (h->storage) {
    case LBM_STORE_F64: k_mass_partial<LBM_STORE_F64><<<nb, 256, 0, S(stream)>>>(p, part); break;
    case LBM_STORE_F32: k_mass_partial<LBM_STORE_F32><<<nb, 256, 0, S(stream)>>>(p, part); break;
    case LBM_STORE_F16: k_mass_partial<LBM_STORE_F16><<<nb, 256, 0, S(stream)>>>(p, part); break;
    case LBM_STORE_BF16: k_mass_partial<LBM_STORE_BF16><<<nb, 256, 0, S(stream)>>>(p, part); break;
    case LBM_STORE_Q16: k_mass_partial<LBM_STORE_Q16><<<nb, 256, 0, S(stream)>>>(p, part); break;
    case LBM_STORE_F16E8: k_mass_partial<LBM_STORE_F16E8><<<nb, 256, 0, S(stream)>>>(p, part); break;
    default: k_mass_partial<LBM_STORE_F16E><<<nb, 256, 0, S(stream)>>>(p, part); break;
    }
    LBM_LAUNCH_CHECK();
    k_sum_serial<<<1, 64, 0, S(stream)>>>(h->d_part, nb, mass_base(h->storage, h->N),
                                          h->d_part + nb);
    LBM_LAUNCH_CHECK();
    LBM_HIP_TRY(hipMemcpyAsync(out, h->d_part + nb, sizeof(double), hipMemcpyDeviceToHost,
                               S(stream)));
    LBM_HIP_TRY(hipStreamSynchronize(S(stream)));
    return LBM_OK;
}

// ================================================================= batch ===
struct lbm_batch {
    std::vector<lbm_fused *> cases;
    std::vector<double *> prof;          // 5 per case
    Params *d_params = nullptr;          // [n_cases][2]
    int2 *d_work = nullptr;
    int n_work = 0;
    double *d_amp = nullptr;             // [ring][n_cases][5]
    int ring = 0, storage = 0, cpt = 1, grid = 0;
    bool vec4 = false;
    int *d_base = nullptr;
    hipStream_t cap_stream = nullptr;
    std::map<GraphKey, hipGraphExec_t> graphs;
};

namespace {

void batch_free(lbm_batch *b)
{
    for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
    for (lbm_fused *h : b->cases) lbm_fused_destroy(h);
    for (double *p : b->prof)
        if (p) (void)hipFree(p);
    void *ptrs[] = {b->d_params, b->d_work, b->d_amp, b->d_base};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
}

template <int S, int CPT, bool V4>
void launch_batch_s(const lbm_batch *b, int k, int parity, hipStream_t s)
{
    k_batch_step<S, CPT, V4><<<b->grid, 256, 0, s>>>(b->d_params, b->d_work, b->n_work, k,
                                                      b->d_base, parity);
}

template <int S> void launch_batch_t(const lbm_batch *b, int k, int parity, hipStream_t s)
{
    if constexpr (S == LBM_STORE_F16 || S == LBM_STORE_BF16) {
        if (b->vec4) {
            switch (b->cpt) {
            case 1: launch_batch_s<S, 1, true>(b, k, parity, s); return;
            case 2: launch_batch_s<S, 2, true>(b, k, parity, s); return;
            default: launch_batch_s<S, 4, true>(b, k, parity, s); return;
            }
        }
    }
    switch (b->cpt) {
    case 1: launch_batch_s<S, 1, false>(b, k, parity, s); break;
    case 2: launch_batch_s<S, 2, false>(b, k, parity, s); break;
    case 4: launch_batch_s<S, 4, false>(b, k, parity, s); break;
    default: launch_batch_s<S, 8, false>(b, k, parity, s); break;
    }
}

void launch_batch(const lbm_batch *b, int k, int parity, hipStream_t s)
{
    switch (b->storage) {
    case LBM_STORE_F64: launch_batch_t<LBM_STORE_F64>(b, k, parity, s); break;
    case LBM_STORE_F32: launch_batch_t<LBM_STORE_F32>(b, k, parity, s); break;
    case LBM_STORE_F16: launch_batch_t<LBM_STORE_F16>(b, k, parity, s); break;
    case LBM_STORE_BF16: launch_batch_t<LBM_STORE_BF16>(b, k, parity, s); break;
    case LBM_STORE_Q16: launch_batch_t<LBM_STORE_Q16>(b, k, parity, s); break;
    case LBM_STORE_F16E8: launch_batch_t<LBM_STORE_F16E8>(b, k, parity, s); break;
    default: launch_batch_t<LBM_STORE_F16E>(b, k, parity, s); break;
    }
}

}  // namespace

extern "C" int lbm_batch_create(const lbm_batch_desc *d, lbm_batch **out)
{
    if (!d || !out || d->n_cases <= 0 || !d->cases || d->ring < 2)
        return set_error(LBM_ERR_ARG, "lbm_batch_create: bad argument");
    *out = nullptr;
    auto b = std::make_unique<lbm_batch>();
    b->ring = d->ring;
    b->storage = d->storage;
    const int n = d->n_cases;
    // one vector width for every case: four nodes per work-item only if all allow it
    int vw = d->vector_width == 4 ? 4 : 1;
    if (vw == 4 && (d->storage == LBM_STORE_F16 || d->storage == LBM_STORE_BF16)) {
        const int cpt = d->cells_per_thread >= 4 || d->cells_per_thread == 0 ? 4 : d->cells_per_thread;
        for (int c = 0; c < n; ++c) {
            const lbm_case_desc &cd = d->cases[c];
            if (cd.ny % 4 || ((long long)cd.nx * cd.ny) % (1024LL * cpt)) vw = 1;
        }
    }
    if (hipMalloc((void **)&b->d_amp, sizeof(double) * (size_t)b->ring * n * 5) != hipSuccess ||
        hipMemset(b->d_amp, 0, sizeof(double) * (size_t)b->ring * n * 5) != hipSuccess) {
        batch_free(b.get());
        return set_error(LBM_ERR_HIP, "batch amplitude table");
    }
    std::vector<Params> params;
    std::vector<int2> work, pwork;
    for (int c = 0; c < n; ++c) {
        const lbm_case_desc &cd = d->cases[c];
        lbm_fused_desc fd{};
        fd.nx = cd.nx;
        fd.ny = cd.ny;
        fd.storage = d->storage;
        fd.om_p = cd.om_p;
        fd.om_m = cd.om_m;
        for (int i = 0; i < 4; ++i) fd.bc[i] = cd.bc[i];
        fd.corners = 1;
        fd.obstacles_first = 1;
        fd.n_links = cd.n_links;
        fd.links_host = cd.links_host;
        fd.ibb_host = cd.ibb_host;
        fd.link_ibb_host = cd.link_ibb_host;
        fd.n_obs = cd.n_obs;
        fd.obs_offsets_host = cd.obs_offsets_host;
        fd.ring = d->ring;
        fd.cells_per_thread = d->cells_per_thread;
        fd.vector_width = vw;
        fd.q16_range_host = d->q16_range_host;
        lbm_fused *h = nullptr;
        if (int e = lbm_fused_create(&fd, &h)) {
            batch_free(b.get());
            return set_error(e, "case " + std::to_string(c) + ": " + g_last_error);
        }
        b->cases.push_back(h);
        if (c == 0) {
            b->cpt = h->cpt;
            b->vec4 = h->vec4;
        } else if (h->cpt != b->cpt || h->vec4 != b->vec4) {
            batch_free(b.get());
            return set_error(LBM_ERR_ARG, "cases resolve to different tile shapes");
        }
        // wall profiles: u_bot, u_left, u_top, u_right ([2][len]) and rho_right ([ny])
        const int len[5] = {2 * cd.nx, 2 * cd.ny, 2 * cd.nx, 2 * cd.ny, cd.ny};
        for (int sd = 0; sd < 5; ++sd) {
            double *dp = nullptr;
            if (hipMalloc((void **)&dp, sizeof(double) * len[sd]) != hipSuccess) {
                batch_free(b.get());
                return set_error(LBM_ERR_HIP, "profile allocation");
            }
            b->prof.push_back(dp);
            std::vector<double> hv(len[sd], 0.0);
            if (cd.profile_host[sd]) std::memcpy(hv.data(), cd.profile_host[sd], sizeof(double) * len[sd]);
            if (hipMemcpy(dp, hv.data(), sizeof(double) * len[sd], hipMemcpyHostToDevice) != hipSuccess) {
                batch_free(b.get());
                return set_error(LBM_ERR_HIP, "profile upload");
            }
            h->prof[sd] = dp;
        }
        h->amp = b->d_amp + (size_t)c * 5;
        h->amp_stride = n * 5;
        for (int parity = 0; parity < 2; ++parity) {
            Params p = base_params(h);
            p.src = h->buf[parity ^ 1];   // iteration of this parity reads the other buffer
            p.dst = h->buf[parity];
            p.rho_out = nullptr;
            p.u_out = nullptr;
            params.push_back(p);
        }
        for (int k = 0; k < h->n_tiles; ++k) work.push_back(make_int2(c, k));
        const int pblocks = (h->n_perim_wg + 255) / 256;
        for (int k = 0; k < pblocks; ++k) pwork.push_back(make_int2(c, h->n_tiles + k));
    }
    // every case's perimeter blocks first: their node-by-node chains then run
    // beside the tiles instead of after them
    work.insert(work.begin(), pwork.begin(), pwork.end());
    // then blocks of 16 consecutive items on one XCD: with one work-group per
    // item, work-group g (on XCD g mod 8) takes item ((i >> 4) * 8 + g mod 8)
    // * 16 + (i mod 16), i = g / 8; the tail of fewer than 8 blocks keeps its
    // order (256-cavity ensemble:
    // FP32 66.6 -> 72.0 GLUPS, FP64 36.6 -> 38.7; blocks of 4: 70.4, of 8:
    // 71.3, of 32: 72.4 / 38.4, of 64: 70.9)
    {
        const int nw = (int)work.size();
        std::vector<int2> order(work);
        for (int g = 0; g < nw; ++g) order[g] = work[xcd_block_unit(g, nw, 4)];
        work.swap(order);
    }
    b->n_work = (int)work.size();
    if (hipMalloc((void **)&b->d_params, sizeof(Params) * params.size()) != hipSuccess ||
        hipMalloc((void **)&b->d_work, sizeof(int2) * work.size()) != hipSuccess ||
        hipMalloc((void **)&b->d_base, sizeof(int)) != hipSuccess ||
        hipMemcpy(b->d_params, params.data(), sizeof(Params) * params.size(),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_work, work.data(), sizeof(int2) * work.size(), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking) != hipSuccess) {
        batch_free(b.get());
        return set_error(LBM_ERR_HIP, "batch tables");
    }
    b->grid = d->grid > 0 ? std::min(d->grid, b->n_work) : b->n_work;
    *out = b.release();
    return LBM_OK;
}

extern "C" void lbm_batch_destroy(lbm_batch *b)
{
    if (!b) return;
    (void)hipDeviceSynchronize();
    batch_free(b);
    delete b;
}

extern "C" int lbm_batch_case(lbm_batch *b, int c, lbm_fused **out)
{
    if (!b || !out || c < 0 || c >= (int)b->cases.size())
        return set_error(LBM_ERR_ARG, "lbm_batch_case: bad argument");
    *out = b->cases[c];
    return LBM_OK;
}

extern "C" int lbm_batch_info(const lbm_batch *b, int32_t *out4)
{
    if (!b || !out4) return set_error(LBM_ERR_ARG, "lbm_batch_info: bad argument");
    out4[0] = b->n_work;
    out4[1] = b->grid;
    out4[2] = b->cpt;
    out4[3] = b->vec4 ? 4 : 1;
    return LBM_OK;
}

extern "C" int lbm_batch_set_amplitudes(lbm_batch *b, int64_t row0, int nrows,
                                        const double *amp_host, void *stream)
{
    if (!b || nrows < 0 || nrows > b->ring || (nrows && !amp_host))
        return set_error(LBM_ERR_ARG, "lbm_batch_set_amplitudes: bad argument");
    LBM_HIP_TRY(hipStreamSynchronize(S(stream)));
    const size_t per_row = b->cases.size() * 5;
    int r = 0;
    while (r < nrows) {
        const int row = (int)(((row0 + r) % b->ring + b->ring) % b->ring);
        const int len = std::min(nrows - r, b->ring - row);
        LBM_HIP_TRY(hipMemcpy(b->d_amp + (size_t)row * per_row, amp_host + (size_t)r * per_row,
                              sizeof(double) * per_row * len, hipMemcpyHostToDevice));
        r += len;
    }
    return LBM_OK;
}

extern "C" int lbm_batch_init(lbm_batch *b, void *stream)
{
    if (!b) return set_error(LBM_ERR_ARG, "lbm_batch_init: bad argument");
    for (lbm_fused *h : b->cases)
        if (int e = lbm_init(h, nullptr, nullptr, nullptr, stream)) return e;
    return LBM_OK;
}

namespace {
int batch_graph(lbm_batch *b, int64_t it0, int n, hipGraphExec_t *out)
{
    const GraphKey key{n, (int)(it0 & 1)};
    auto itg = b->graphs.find(key);
    if (itg == b->graphs.end()) {
        hipGraph_t graph;
        LBM_HIP_TRY(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < n; ++k) launch_batch(b, k, (int)((it0 + k) & 1), b->cap_stream);
        hipError_t ec = hipStreamEndCapture(b->cap_stream, &graph);
        if (ec != hipSuccess)
            return set_error(LBM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
        hipGraphExec_t exec;
        ec = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ec != hipSuccess)
            return set_error(LBM_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ec));
        itg = b->graphs.emplace(key, exec).first;
    }
    *out = itg->second;
    return LBM_OK;
}
}  // namespace

extern "C" int lbm_batch_steps(lbm_batch *b, int64_t it0, int n, void *stream)
{
    if (!b || it0 < 1 || n < 0) return set_error(LBM_ERR_ARG, "lbm_batch_steps: bad argument");
    if (n == 0) return LBM_OK;
    hipGraphExec_t exec;
    if (int e = batch_graph(b, it0, n, &exec)) return e;
    const int base = (int)(((it0 - 1) % b->ring + b->ring) % b->ring);
    k_set_int<<<1, 64, 0, S(stream)>>>(b->d_base, base);
    LBM_LAUNCH_CHECK();
    LBM_HIP_TRY(hipGraphLaunch(exec, S(stream)));
    return LBM_OK;
}

extern "C" int lbm_batch_steps_prepare(lbm_batch *b, int64_t it0, int n, void *stream)
{
    if (!b || it0 < 1 || n < 0)
        return set_error(LBM_ERR_ARG, "lbm_batch_steps_prepare: bad argument");
    if (n == 0) return LBM_OK;
    hipGraphExec_t exec;
    if (int e = batch_graph(b, it0, n, &exec)) return e;
    LBM_HIP_TRY(hipGraphUpload(exec, S(stream)));
    return LBM_OK;
}
