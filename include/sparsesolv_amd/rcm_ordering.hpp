// rcm_ordering.hpp -- reverse Cuthill-McKee ordering (host C++17), the
// first stage of the ABMC ordering when abmc_use_rcm is set
// (core/rcm_ordering.hpp:37-141, ic_preconditioner.hpp:97-103).
//
// The steps are the reference's, kept alike so that the ordering, and with
// it the ABMC schedule and the factor, are the same:
//   * start at the first row of least degree (row length, diagonal included);
//   * two breadth-first passes, each restarting from the first row reached
//     at the greatest distance of the previous one (a pseudo-peripheral row);
//   * Cuthill-McKee order: breadth-first from that row, each row's unvisited
//     neighbours (in column order) sorted by ascending degree with std::sort
//     -- not a stable sort, so rows of equal degree land where that sort puts
//     them, which the same call on the same sequence reproduces -- and, for
//     rows not reached, from each unvisited row in index order;
//   * reversed.
#pragma once

#include <algorithm>
#include <queue>
#include <vector>

#include "abmc_ordering.hpp"

namespace sparsesolv_amd {

// ordering[old] = new, reverse_ordering[new] = old
inline void rcm_ordering(const index_t* rp, const index_t* ci, index_t n,
                         std::vector<index_t>& ordering, std::vector<index_t>& reverse_ordering)
{
    ordering.assign(size_t(std::max<index_t>(n, 0)), 0);
    reverse_ordering.assign(size_t(std::max<index_t>(n, 0)), 0);
    if (n <= 0) return;
    std::vector<index_t> degree(static_cast<size_t>(n));
    for (index_t i = 0; i < n; ++i) degree[size_t(i)] = rp[i + 1] - rp[i];

    index_t start = 0;
    for (index_t i = 1; i < n; ++i)
        if (degree[size_t(i)] < degree[size_t(start)]) start = i;

    std::vector<index_t> dist(static_cast<size_t>(n));
    for (int pass = 0; pass < 2; ++pass) {
        std::fill(dist.begin(), dist.end(), -1);
        std::queue<index_t> q;
        q.push(start);
        dist[size_t(start)] = 0;
        index_t far = start, far_d = 0;
        while (!q.empty()) {
            const index_t u = q.front();
            q.pop();
            for (index_t k = rp[u]; k < rp[u + 1]; ++k) {
                const index_t v = ci[k];
                if (dist[size_t(v)] >= 0) continue;
                dist[size_t(v)] = dist[size_t(u)] + 1;
                q.push(v);
                if (dist[size_t(v)] > far_d) {
                    far_d = dist[size_t(v)];
                    far = v;
                }
            }
        }
        start = far;
    }

    std::vector<char> seen(static_cast<size_t>(n), 0);
    std::vector<index_t> cm;
    cm.reserve(static_cast<size_t>(n));
    std::vector<index_t> next;
    for (index_t s = 0; s < n; ++s) {
        const index_t seed = s == 0 ? start : s;
        if (seen[size_t(seed)]) continue;
        std::queue<index_t> q;
        q.push(seed);
        seen[size_t(seed)] = 1;
        cm.push_back(seed);
        while (!q.empty()) {
            const index_t u = q.front();
            q.pop();
            next.clear();
            for (index_t k = rp[u]; k < rp[u + 1]; ++k) {
                const index_t v = ci[k];
                if (!seen[size_t(v)]) {
                    next.push_back(v);
                    seen[size_t(v)] = 1;
                }
            }
            std::sort(next.begin(), next.end(),
                      [&](index_t a, index_t b) { return degree[size_t(a)] < degree[size_t(b)]; });
            for (index_t v : next) {
                cm.push_back(v);
                q.push(v);
            }
        }
    }
    for (index_t i = 0; i < n; ++i) {
        const index_t old = cm[size_t(n - 1 - i)];
        reverse_ordering[size_t(i)] = old;
        ordering[size_t(old)] = i;
    }
}

}  // namespace sparsesolv_amd
