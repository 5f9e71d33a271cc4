// Reads "m cap s_0 ... s_{m-1}" from stdin and prints the batch boundaries
// batch_cuts gives for those row sizes, then the largest batch total.
#include <iostream>
#include <vector>

#include "batching.h"

int main() {
    int64_t m = 0, cap = 0;
    std::cin >> m >> cap;
    std::vector<int64_t> prefix(m + 1, 0);
    for (int64_t r = 0; r < m; r++) {
        int64_t x = 0;
        std::cin >> x;
        prefix[r + 1] = prefix[r] + x;
    }
    int64_t largest = -1;
    const std::vector<int64_t> cuts = pggc::batch_cuts(prefix, cap, &largest);
    for (int64_t c : cuts) std::cout << c << " ";
    std::cout << "\n" << largest << "\n";
    return 0;
}
