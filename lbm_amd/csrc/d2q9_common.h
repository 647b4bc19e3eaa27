// d2q9_common.h -- shared definitions of the gfx950 D2Q9 kernels.
//
// Velocity set, weights and bounce-back table are the D2Q9 tables of the
// reference (lbm/src/core/lattice.py:135-152):
//   c  = (0,0) (1,0) (-1,0) (0,1) (0,-1) (1,1) (-1,-1) (-1,1) (1,-1)
//   w  = 4/9, 1/9 x4, 1/36 x4
//   ns = 0 2 1 4 3 6 5 8 7
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

namespace lbm {

constexpr int CX[9] = {0, 1, -1, 0, 0, 1, -1, -1, 1};
constexpr int CY[9] = {0, 0, 0, 1, -1, 1, -1, 1, -1};
constexpr int NS[9] = {0, 2, 1, 4, 3, 6, 5, 8, 7};

template <typename T>
__host__ __device__ constexpr T weight(int q)
{
    return q == 0 ? T(4.0 / 9.0) : (q <= 4 ? T(1.0 / 9.0) : T(1.0 / 36.0));
}

// Work-group b of nb (work-groups are dealt round-robin over the 8 XCDs, so b
// runs on XCD b mod 8) -> the unit of work it takes when units go to the XCDs
// in blocks of 2^bs consecutive ones: block (i >> bs) * 8 + b mod 8, unit
// i mod 2^bs of it, i = b / 8; the tail of fewer than 8 blocks keeps its
// order.  bs = 0: b itself.
__host__ __device__ constexpr int xcd_block_unit(int b, int nb, int bs)
{
    return (bs > 0 && b < nb - nb % (8 << bs))
               ? (((((b >> 3) >> bs) << 3) + (b & 7)) << bs) + ((b >> 3) & ((1 << bs) - 1))
               : b;
}

// Error reporting of the C ABI (per calling thread).
int set_error(int code, const std::string &msg);

#define LBM_HIP_TRY(expr)                                                        \
    do {                                                                         \
        hipError_t e_ = (expr);                                                  \
        if (e_ != hipSuccess)                                                    \
            return ::lbm::set_error(-2, std::string(#expr) + ": " +              \
                                            hipGetErrorString(e_));              \
    } while (0)

#define LBM_LAUNCH_CHECK()                                                       \
    do {                                                                         \
        hipError_t e_ = hipGetLastError();                                       \
        if (e_ != hipSuccess)                                                    \
            return ::lbm::set_error(-2, std::string("kernel launch: ") +         \
                                            hipGetErrorString(e_));              \
    } while (0)

}  // namespace lbm
