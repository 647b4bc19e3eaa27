// ThreadSanitizer driver for the library's host code that computes moving
// averages over a chunk of iterations (lbm_mv_avg_chunk, several threads with
// a barrier between its levels): built from lbm_amd/csrc/d2q9_observe.cpp
// with -fsanitize=thread by tests/test_observe_sanitize.py and run as a
// child process.  A chunk large enough to take the threaded path is computed
// twice, on several threads (argument: thread count) and on one, and the two
// results must agree bit for bit; a data race makes the sanitizer fail the
// run.  Exit status 0 with "ok" on stdout.
#include "../include/lbm_d2q9.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

static int run(const std::vector<double> &hist, int64_t it0, int k, const char *threads,
               std::vector<double> &a1, std::vector<double> &a2, std::vector<double> &a3,
               std::vector<double> &obs)
{
    setenv("OMP_NUM_THREADS", threads, 1);
    const size_t n = hist.size();
    a1.assign(n, 0.0);
    a2.assign(n, 0.0);
    a3.assign(n, 0.0);
    obs.assign((size_t)k, 0.0);
    // averages of iterations 1 .. it0 first (one thread: small per call)
    for (int64_t it = 0; it < it0; it += 500) {
        const int kk = (int)std::min<int64_t>(500, it0 - it);
        std::vector<double> tmp((size_t)kk);
        if (lbm_mv_avg_chunk(hist.data(), a1.data(), a2.data(), a3.data(), it, kk, tmp.data()))
            return 1;
    }
    return lbm_mv_avg_chunk(hist.data(), a1.data(), a2.data(), a3.data(), it0, k, obs.data());
}

int main(int argc, char **argv)
{
    const char *threads = argc > 1 ? argv[1] : "8";
    const int64_t it0 = 6000;
    const int k = 1000;                  // k * (it0 + k) = 7e6: the threaded path
    std::vector<double> hist((size_t)(it0 + k + 2));
    for (size_t i = 0; i < hist.size(); ++i)
        hist[i] = 3.2 + 0.3 * std::sin(0.013 * (double)i) - 2.0 * std::exp(-(double)i / 50.0);
    std::vector<double> a1, a2, a3, o, b1, b2, b3, p;
    if (run(hist, it0, k, threads, a1, a2, a3, o)) return 2;
    if (run(hist, it0, k, "1", b1, b2, b3, p)) return 3;
    if (a1 != b1 || a2 != b2 || a3 != b3 || o != p) return 4;
    std::printf("ok %s threads, last average %.17g\n", threads, o[(size_t)k - 1]);
    return 0;
}
