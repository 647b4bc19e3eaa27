// d2q9_observe.cpp -- the moving averages of an observable history
// (lbm/src/utils/buff.py:51-77, restated in lbm_amd/buff.py) for a chunk of
// iterations at once, host code.
//
// buff.mv_avg() averages the last fifth of a history three times over with
// np.sum; its cost grows with the iteration count (O(it) per iteration).  This
// file computes the same numbers, bit for bit, for K consecutive iterations:
//
// * np.sum of a contiguous float64 slice is a sequential sum, starting from 0,
//   of NumPy's pairwise sums of consecutive blocks of 8192 elements (the
//   reduction buffer size); a pairwise sum of n elements is a plain loop for
//   n < 8, eight interleaved accumulators combined as ((0+1)+(2+3))+((4+5)+
//   (6+7)) plus a sequential tail for n <= 128, and the sum of the two halves
//   split at n/2 rounded down to a multiple of 8 above that.  tests/
//   test_observe.py checks this against np.sum itself.
// * Each of the four sums of an iteration k (it_s = floor(4k/5), window
//   [it_s, k)) reads only entries written before iteration k - 1, so the
//   chunk is computed level by level: the K window sums over the history,
//   then over the first averages (now known for the whole chunk), and so on,
//   each level parallel over the K iterations.
#include "../../include/lbm_d2q9.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <thread>
#include <vector>

namespace {

double pairwise(const double *a, int64_t n)
{
    if (n < 8) {
        double r = 0.0;
        for (int64_t i = 0; i < n; ++i) r += a[i];
        return r;
    }
    if (n <= 128) {
        double r[8];
        for (int j = 0; j < 8; ++j) r[j] = a[j];
        int64_t i = 8;
        for (; i < n - (n % 8); i += 8)
            for (int j = 0; j < 8; ++j) r[j] += a[i + j];
        double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < n; ++i) res += a[i];
        return res;
    }
    int64_t n2 = n / 2;
    n2 -= n2 % 8;
    return pairwise(a, n2) + pairwise(a + n2, n - n2);
}

// np.sum(a[0:n]) for a contiguous float64 array
double np_sum(const double *a, int64_t n)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; i += 8192) s = s + pairwise(a + i, n - i < 8192 ? n - i : 8192);
    return s;
}

// buff.p_avg(b, i, j) = np.sum(b[i:j]) / float(j - i + 1)
double p_avg(const double *b, int64_t i, int64_t j)
{
    return np_sum(b + i, j > i ? j - i : 0) / (double)(j - i + 1);
}

// Worker threads: OMP_NUM_THREADS if set, else the hardware threads, at most 16.
int worker_count()
{
    int n = (int)std::thread::hardware_concurrency();
    if (const char *e = std::getenv("OMP_NUM_THREADS")) n = std::atoi(e);
    return std::max(1, std::min(n, 16));
}

}  // namespace

extern "C" int lbm_mv_avg_chunk(const double *hist, double *avg1, double *avg2, double *avg3,
                                int64_t it0, int k, double *obs_out)
{
    if (it0 < 0 || k < 0 || (k && (!hist || !avg1 || !avg2 || !avg3 || !obs_out)))
        return LBM_ERR_ARG;
    // iteration it = it0 + 1 + t (t < k): the history holds entries 0 .. it+1,
    // level l writes entry it + 1 of average l+1 (the last level: obs_out[t])
    double *lev[4] = {avg1, avg2, avg3, nullptr};
    const double *src[4] = {hist, avg1, avg2, avg3};
    auto level = [&](int l, int t0, int step) {
        for (int t = t0; t < k; t += step) {
            const int64_t it = it0 + 1 + t;
            const int64_t it_s = (int64_t)std::floor((double)(it * 4) / 5.0);
            const double v = p_avg(src[l], it_s, it);
            if (lev[l]) lev[l][it + 1] = v;
            else obs_out[t] = v;
        }
    };
    // window sums of this chunk: ~ k * it / 5 additions per level
    const int nt = ((double)k * (double)(it0 + k) > 4.0e6) ? std::min(worker_count(), k) : 1;
    if (nt == 1) {
        for (int l = 0; l < 4; ++l) level(l, 0, 1);
        return LBM_OK;
    }
    // one pass per level, split over the threads, a barrier between levels
    std::atomic<int> arrived{0};
    auto work = [&](int tid) {
        for (int l = 0; l < 4; ++l) {
            level(l, tid, nt);
            arrived.fetch_add(1);
            while (arrived.load() < nt * (l + 1)) std::this_thread::yield();
        }
    };
    std::vector<std::thread> pool;
    for (int i = 1; i < nt; ++i) pool.emplace_back(work, i);
    work(0);
    for (auto &th : pool) th.join();
    return LBM_OK;
}
