// d2q9_probe.hip -- HBM copy probe: the bandwidth ceiling the D2Q9 step is
// measured against, and the reference kernel of known traffic that scales
// the FETCH_SIZE / WRITE_SIZE counters.  Not a reference function.
#include "d2q9_common.h"
#include "../../include/lbm_d2q9.h"

namespace lbm {
namespace {

// Each work-item moves UNROLL 16-byte vectors per pass, all loads first.
constexpr int UNROLL = 4;

__global__ __launch_bounds__(256) void k_copy16(uint4 *__restrict__ dst,
                                                const uint4 *__restrict__ src, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * 256 * UNROLL;
    for (int64_t i = (int64_t)blockIdx.x * 256 * UNROLL + threadIdx.x; i < n; i += stride) {
        uint4 v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u)
            if (i + u * 256 < n) v[u] = src[i + u * 256];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u)
            if (i + u * 256 < n) dst[i + u * 256] = v[u];
    }
}

// The same with non-temporal loads and stores (no cache allocation).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_copy16_nt(u32x4 *__restrict__ dst,
                                                   const u32x4 *__restrict__ src, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * 256 * UNROLL;
    for (int64_t i = (int64_t)blockIdx.x * 256 * UNROLL + threadIdx.x; i < n; i += stride) {
        u32x4 v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u)
            if (i + u * 256 < n) v[u] = __builtin_nontemporal_load(src + i + u * 256);
#pragma unroll
        for (int u = 0; u < UNROLL; ++u)
            if (i + u * 256 < n) __builtin_nontemporal_store(v[u], dst + i + u * 256);
    }
}

}  // namespace
}  // namespace lbm

extern "C" int lbm_copy_probe(void *dst, const void *src, int64_t bytes, int blocks,
                              void *stream)
{
    if (!dst || !src || bytes <= 0 || (bytes & 15) || blocks <= 0)
        return lbm::set_error(LBM_ERR_ARG, "lbm_copy_probe: bad argument");
    const int64_t n = bytes / 16;
    lbm::k_copy16<<<blocks, 256, 0, reinterpret_cast<hipStream_t>(stream)>>>(
        static_cast<uint4 *>(dst), static_cast<const uint4 *>(src), n);
    LBM_LAUNCH_CHECK();
    return LBM_OK;
}

extern "C" int lbm_copy_probe_nt(void *dst, const void *src, int64_t bytes, int blocks,
                                 void *stream)
{
    if (!dst || !src || bytes <= 0 || (bytes & 15) || blocks <= 0)
        return lbm::set_error(LBM_ERR_ARG, "lbm_copy_probe_nt: bad argument");
    const int64_t n = bytes / 16;
    lbm::k_copy16_nt<<<blocks, 256, 0, reinterpret_cast<hipStream_t>(stream)>>>(
        static_cast<lbm::u32x4 *>(dst), static_cast<const lbm::u32x4 *>(src), n);
    LBM_LAUNCH_CHECK();
    return LBM_OK;
}
