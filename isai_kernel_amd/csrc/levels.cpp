// levels.cpp -- level sets of the block triangular dependency graph (host).
//
// level(i) = 1 + max level(k) over stored blocks (i,k) with k < i (lower) or
// k > i (upper, counted from the last row); 0 when row i has no such block.
// The level-ordered row list (stable counting sort: level ascending, row index
// ascending inside a level) is the dispatch order of the ILU(0) factorization.
#include <cstring>
#include <vector>

#include "bsrilu.h"
#include "host_common.hpp"
#include "levels.hpp"

namespace isai {

int compute_levels(int nb, const int *rp, const int *ci, bool lower, int *level, int *nlevels) {
    int nl = 0;
    if (lower) {
        for (int i = 0; i < nb; ++i) {
            int l = 0;
            for (int p = rp[i]; p < rp[i + 1]; ++p) {
                const int k = ci[p];
                if (k < i && level[k] + 1 > l) l = level[k] + 1;
            }
            level[i] = l;
            if (l + 1 > nl) nl = l + 1;
        }
    } else {
        for (int i = nb - 1; i >= 0; --i) {
            int l = 0;
            for (int p = rp[i]; p < rp[i + 1]; ++p) {
                const int k = ci[p];
                if (k > i && level[k] + 1 > l) l = level[k] + 1;
            }
            level[i] = l;
            if (l + 1 > nl) nl = l + 1;
        }
    }
    *nlevels = nl;
    return ISAI_OK;
}

void level_order(int nb, const int *level, int nlevels, int *order, int *level_ptr) {
    std::vector<int> cnt(nlevels + 1, 0);
    for (int i = 0; i < nb; ++i) cnt[level[i] + 1]++;
    for (int l = 0; l < nlevels; ++l) cnt[l + 1] += cnt[l];
    if (level_ptr) std::memcpy(level_ptr, cnt.data(), sizeof(int) * (nlevels + 1));
    for (int i = 0; i < nb; ++i) order[cnt[level[i]]++] = i;
}

}  // namespace isai

extern "C" int isai_bsr_levels(int nb, const int *Arow, const int *Acol, int lower, int *level,
                               int *nlevels) {
    ISAI_HOST_REQUIRE(nb >= 0 && Arow && Acol && level && nlevels, "isai_bsr_levels: bad argument");
    return isai::compute_levels(nb, Arow, Acol, lower != 0, level, nlevels);
}
