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
// Float forms: every corner's assignments are evaluated from the incoming
// values (no corner reads a population it writes) and the result is selected
// per population, pairwise.  With a switch the compiler turned the zeroed
// pair of diagonals into a store through an index chosen at run time, which
// placed the population array in scratch memory in every kernel inlining this
// routine.  FP64 keeps the switch: there the selected form raised the register
// peak of the 7-waves-per-SIMD step kernel (72 VGPRs) into heavy spilling,
// while the switch's 48-byte scratch array costs nothing on the node path.
template <typename T> __device__ __forceinline__ void zh_corner(T *f, int which, T r, T ux, T uy)
{
    if constexpr (sizeof(T) == 8) {
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
    } else {
        const bool w0 = which == 0, w1 = which == 1, w2 = which == 2, w3 = !(w0 || w1 || w2);
        const T n1 = (w0 || w1) ? f[2] + T(K1) * r * ux : f[1];                     // BL, TL
        const T n2 = (w2 || w3) ? f[1] - T(K1) * r * ux : f[2];                     // TR, BR
        f[1] = n1;
        f[2] = n2;
        const T n3 = (w0 || w3) ? f[4] + T(K1) * r * uy : f[3];                     // BL, BR
        const T n4 = (w1 || w2) ? f[3] - T(K1) * r * uy : f[4];                     // TL, TR
        f[3] = n3;
        f[4] = n4;
        const T n5 = w0 ? f[6] + T(K2) * r * ux + T(K2) * r * uy : ((w1 || w3) ? T(0.0) : f[5]);
        const T n6 = w2 ? f[5] - T(K2) * r * ux - T(K2) * r * uy : ((w1 || w3) ? T(0.0) : f[6]);
        f[5] = n5;
        f[6] = n6;
        const T n7 = w3 ? f[8] - T(K2) * r * ux + T(K2) * r * uy : ((w0 || w2) ? T(0.0) : f[7]);
        const T n8 = w1 ? f[7] + T(K2) * r * ux - T(K2) * r * uy : ((w0 || w2) ? T(0.0) : f[8]);
        f[7] = n7;
        f[8] = n8;
    }
    T s = r;
#pragma unroll
    for (int q = 1; q < 9; ++q) s = s - f[q];
    f[0] = s;
}


}  // namespace lbm
