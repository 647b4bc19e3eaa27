// Host check of lbm::xcd_block_unit (lbm_amd/csrc/d2q9_common.h), the
// work-group -> unit map the step kernels, the batch work list and the
// adjoint kernel use to deal blocks of consecutive units to the XCDs (test
// infrastructure, built by tests/test_block_map.py).  For every size and
// block shift: a permutation of [0, nb); work-groups on one XCD (b mod 8)
// take whole blocks in order; the tail keeps its order; bs = 0 is the
// identity.  Exit status 0 with "ok" on stdout.
#include "../lbm_amd/csrc/d2q9_common.h"

#include <cstdio>
#include <vector>

namespace lbm {
int set_error(int code, const std::string &) { return code; }
}

int main()
{
    const int sizes[] = {1, 7, 8, 9, 63, 64, 65, 255, 256, 1000, 4096, 12345, 32768, 65536, 107248};
    long checks = 0;
    for (int nb : sizes) {
        for (int bs = 0; bs <= 6; ++bs) {
            std::vector<int> seen(nb, 0);
            const int span = 8 << bs, full = nb - nb % span;
            for (int b = 0; b < nb; ++b) {
                const int t = lbm::xcd_block_unit(b, nb, bs);
                if (t < 0 || t >= nb || seen[t]++) {
                    std::printf("not a permutation: nb %d bs %d b %d -> %d\n", nb, bs, b, t);
                    return 1;
                }
                if (bs == 0 || b >= full) {
                    if (t != b) {
                        std::printf("identity expected: nb %d bs %d b %d -> %d\n", nb, bs, b, t);
                        return 2;
                    }
                } else {
                    // the unit's block belongs to this work-group's XCD, and the
                    // XCD's i-th work-group takes unit i mod 2^bs of its block
                    const int blk = t >> bs, xcd = b & 7, i = b >> 3;
                    if ((blk & 7) != xcd || (t & ((1 << bs) - 1)) != (i & ((1 << bs) - 1)) ||
                        (blk >> 3) != (i >> bs)) {
                        std::printf("block layout: nb %d bs %d b %d -> %d\n", nb, bs, b, t);
                        return 3;
                    }
                }
                ++checks;
            }
        }
    }
    std::printf("ok %ld\n", checks);
    return 0;
}
