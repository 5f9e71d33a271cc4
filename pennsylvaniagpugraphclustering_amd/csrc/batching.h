// Host-side batching of consecutive rows under a size cap (no GPU code, so
// it is also built and tested on the CPU).
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

namespace pggc {

// prefix[0..m] is the running size of rows 0..m-1 (prefix[0] = 0).  Returns
// the batch boundaries 0 = cuts[0] < cuts[1] < ... < cuts.back() = m: each
// batch [cuts[q], cuts[q+1]) holds at most `cap` in total, except a single
// row larger than `cap`, which is a batch of its own.  *largest receives
// the largest batch total.
inline std::vector<int64_t> batch_cuts(const std::vector<int64_t>& prefix, int64_t cap, int64_t* largest) {
    const int64_t m = (int64_t)prefix.size() - 1;
    std::vector<int64_t> cuts{0};
    int64_t big = 0;
    while (cuts.back() < m) {
        const int64_t r0 = cuts.back();
        int64_t r1 = r0 + 1;
        while (r1 < m && prefix[r1 + 1] - prefix[r0] <= cap) r1++;
        big = std::max(big, prefix[r1] - prefix[r0]);
        cuts.push_back(r1);
    }
    if (largest) *largest = big;
    return cuts;
}

}  // namespace pggc
