// Placement of every grid's chain of passes into the rounds of a batch schedule (host only).
//
// A grid of a batch is advanced by P consecutive passes (P = ceil(steps / T)); a round is a
// set of passes that run together, and rounds run one after the other. A grid's passes must
// sit in P consecutive rounds, but where the chain starts is free: grids are independent.
// Starting every grid in round 0 crowds the first rounds and leaves the last ones nearly
// empty (they then cost a pass's latency for little work). Here each chain gets the start
// that keeps the rounds' work level: grids are placed largest total work first, each at the
// start whose heaviest round after placement is lightest (ties: least work over the window,
// then the earliest start). R = the longest chain, so the round count does not grow.
#ifndef CELLATO_HIP_ROUNDS_HPP
#define CELLATO_HIP_ROUNDS_HPP

#include <algorithm>
#include <deque>
#include <limits>
#include <numeric>
#include <vector>

namespace cellato::hip {

// work[g][r]: work of pass r of grid g; every pass but the last has the same work.
// Returns the start round of each grid; every start + P(g) <= R.
inline std::vector<int> place_chains(const std::vector<std::vector<double>>& work, int R) {
    const int n = static_cast<int>(work.size());
    std::vector<int> start(n, 0);
    std::vector<double> total(n, 0.0);
    for (int g = 0; g < n; ++g) total[g] = std::accumulate(work[g].begin(), work[g].end(), 0.0);
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return total[a] > total[b]; });
    std::vector<double> load(R, 0.0), prefix(R + 1, 0.0);
    for (int g : order) {
        const int P = static_cast<int>(work[g].size());
        if (P == 0 || P > R) continue;
        const double w_full = work[g][0], w_last = work[g][P - 1];
        for (int r = 0; r < R; ++r) prefix[r + 1] = prefix[r] + load[r];
        // sliding maximum of load over rounds [s, s + P - 2] (the full passes)
        std::deque<int> q;
        double best_max = std::numeric_limits<double>::infinity(), best_sum = best_max;
        int best_s = 0;
        int next = 0;     // next round to enter the window
        for (int s = 0; s + P <= R; ++s) {
            double m = load[s + P - 1] + w_last;
            if (P >= 2) {
                while (next <= s + P - 2) {
                    while (!q.empty() && load[q.back()] <= load[next]) q.pop_back();
                    q.push_back(next++);
                }
                while (q.front() < s) q.pop_front();
                m = std::max(m, load[q.front()] + w_full);
            }
            const double sum = prefix[s + P] - prefix[s];
            if (m < best_max || (m == best_max && sum < best_sum)) {
                best_max = m;
                best_sum = sum;
                best_s = s;
            }
        }
        start[g] = best_s;
        for (int r = 0; r < P; ++r) load[best_s + r] += work[g][r];
    }
    return start;
}

} // namespace cellato::hip

#endif
