// d2q9_zouhe.h -- Zou-He wall and corner routines on one node's populations
// held in registers (lbm/src/core/nb.py:121-344), in the reference's
// evaluation order.  T is double (bit-identical to the reference arithmetic
// when compiled without FMA contraction) or float.
#pragma once
#include "d2q9_common.h"

namespace lbm {

constexpr double K1 = 2.0 / 3.0, K2 = 1.0 / 6.0, K3 = 1.0 / 2.0;

// Zou-He walls on a node's populations, nb.py:121-247, same evaluation order.
template <typename T> __device__ __forceinline__ void zh_left(T *f, T ux, T uy, T &r)
{
    r = (f[0] + f[3] + f[4] + T(2.0) * f[2] + T(2.0) * f[6] + T(2.0) * f[7]) / (T(1.0) - ux);
    f[1] = f[2] + T(K1) * r * ux;
    f[5] = f[6] - T(K3) * (f[3] - f[4]) + T(K2) * r * ux + T(K3) * r * uy;
    f[8] = f[7] + T(K3) * (f[3] - f[4]) + T(K2) * r * ux - T(K3) * r * uy;
}
template <typename T> __device__ __forceinline__ void zh_right_tail(T *f, T ux, T uy, T r)
{
    f[2] = f[1] - T(K1) * r * ux;
    f[6] = f[5] + T(K3) * (f[3] - f[4]) - T(K2) * r * ux - T(K3) * r * uy;
    f[7] = f[8] - T(K3) * (f[3] - f[4]) - T(K2) * r * ux + T(K3) * r * uy;
}
template <typename T> __device__ __forceinline__ void zh_right_vel(T *f, T ux, T uy, T &r)
{
    r = (f[0] + f[3] + f[4] + T(2.0) * f[1] + T(2.0) * f[5] + T(2.0) * f[8]) / (T(1.0) + ux);
    zh_right_tail(f, ux, uy, r);
}
template <typename T> __device__ __forceinline__ void zh_right_pres(T *f, T r, T uy, T &ux)
{
    ux = (f[0] + f[3] + f[4] + T(2.0) * f[1] + T(2.0) * f[5] + T(2.0) * f[8]) / r - T(1.0);
    zh_right_tail(f, ux, uy, r);
}
template <typename T> __device__ __forceinline__ void zh_top(T *f, T ux, T uy, T &r)
{
    r = (f[0] + f[1] + f[2] + T(2.0) * f[3] + T(2.0) * f[5] + T(2.0) * f[7]) / (T(1.0) + uy);
    f[4] = f[3] - T(K1) * r * uy;
    f[8] = f[7] - T(K3) * (f[1] - f[2]) + T(K3) * r * ux - T(K2) * r * uy;
    f[6] = f[5] + T(K3) * (f[1] - f[2]) - T(K3) * r * ux - T(K2) * r * uy;
}
template <typename T> __device__ __forceinline__ void zh_bottom(T *f, T ux, T uy, T &r)
{
    r = (f[0] + f[1] + f[2] + T(2.0) * f[4] + T(2.0) * f[6] + T(2.0) * f[8]) / (T(1.0) - uy);
    f[3] = f[4] + T(K1) * r * uy;
    f[5] = f[6] - T(K3) * (f[1] - f[2]) + T(K3) * r * ux + T(K2) * r * uy;
    f[7] = f[8] + T(K3) * (f[1] - f[2]) - T(K3) * r * ux + T(K2) * r * uy;
}

// Zou-He corners, nb.py:251-344.  which: 0 BL, 1 TL, 2 TR, 3 BR.
template <typename T> __device__ __forceinline__ void zh_corner(T *f, int which, T r, T ux, T uy)
{
    switch (which) {
    case 0:
        f[1] = f[2] + T(K1) * r * ux;
        f[3] = f[4] + T(K1) * r * uy;
        f[5] = f[6] + T(K2) * r * ux + T(K2) * r * uy;
        f[7] = T(0.0);
        f[8] = T(0.0);
        break;
    case 1:
        f[1] = f[2] + T(K1) * r * ux;
        f[4] = f[3] - T(K1) * r * uy;
        f[8] = f[7] + T(K2) * r * ux - T(K2) * r * uy;
        f[5] = T(0.0);
        f[6] = T(0.0);
        break;
    case 2:
        f[2] = f[1] - T(K1) * r * ux;
        f[4] = f[3] - T(K1) * r * uy;
        f[6] = f[5] - T(K2) * r * ux - T(K2) * r * uy;
        f[7] = T(0.0);
        f[8] = T(0.0);
        break;
    default:
        f[2] = f[1] - T(K1) * r * ux;
        f[3] = f[4] + T(K1) * r * uy;
        f[7] = f[8] - T(K2) * r * ux + T(K2) * r * uy;
        f[5] = T(0.0);
        f[6] = T(0.0);
        break;
    }
    T s = r;
#pragma unroll
    for (int q = 1; q < 9; ++q) s = s - f[q];
    f[0] = s;
}


}  // namespace lbm
