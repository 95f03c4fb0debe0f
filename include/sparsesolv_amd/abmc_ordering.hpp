// abmc_ordering.hpp -- Algebraic Block Multi-Colour ordering (host C++17).
//
// Produces the same schedule as sparsesolv::ABMCSchedule::build
// (core/abmc_ordering.hpp:83-121) -- identical ordering, colour sets and
// block boundaries -- because the GPU sweeps must follow the reference's
// dependency structure exactly.  The steps that fix the result are kept:
//   * BFS aggregation seeded at the lowest unassigned row, FIFO queue,
//     neighbours marked "queued" so each is enqueued once per block, the
//     block closes at `block_size` rows and queued-but-unassigned rows are
//     released (:147-215); rows left over once ceil(n/B) blocks exist join
//     the last block (:218-226);
//   * block adjacency in first-encounter order, rows of a block visited in
//     insertion order (:235-273);
//   * colour count = max(target, max lower block degree + 1), first-fit
//     colouring against lower-numbered neighbours, empty colours dropped
//     (:288-353);
//   * colour-major, block-ordered, row-sorted renumbering (:365-410).
// The data structures are flat arrays (one allocation per stage) rather
// than vectors of vectors.
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

namespace sparsesolv_amd {

using index_t = std::int32_t;

struct AbmcSchedule {
    std::vector<index_t> color_offsets;   // nc+1, offsets into blocks
    std::vector<index_t> block_offsets;   // nb+1, offsets into new rows
    std::vector<index_t> ordering;        // old row -> new row
    std::vector<index_t> reverse_ordering;// new row -> old row

    index_t num_colors() const { return color_offsets.empty() ? 0 : index_t(color_offsets.size()) - 1; }
    index_t num_blocks() const { return block_offsets.empty() ? 0 : index_t(block_offsets.size()) - 1; }
    bool built() const { return !color_offsets.empty(); }

    void build(const index_t* rp, const index_t* ci, index_t n, int block_size, int target_colors)
    {
        color_offsets.clear(); block_offsets.clear(); ordering.clear(); reverse_ordering.clear();
        if (n <= 0) return;
        if (block_size < 1) block_size = 1;
        if (target_colors < 2) target_colors = 2;

        // ---- aggregation ------------------------------------------------
        const index_t max_blocks = (n + block_size - 1) / block_size;
        std::vector<index_t> owner(n, -1);          // -1 free, -2 queued, >=0 block
        std::vector<index_t> members;               // rows, block-major
        std::vector<index_t> first;                 // block -> offset in members
        std::vector<index_t> fifo(n);
        members.reserve(n);
        first.reserve(max_blocks + 1);
        index_t scan = 0;
        for (index_t blk = 0; blk < max_blocks; ++blk) {
            while (scan < n && owner[scan] >= 0) ++scan;
            if (scan >= n) break;
            const index_t id = index_t(first.size());
            first.push_back(index_t(members.size()));
            index_t head = 0, tail = 0, size = 1;
            owner[scan] = id;
            members.push_back(scan);
            for (index_t k = rp[scan]; k < rp[scan + 1]; ++k)
                if (owner[ci[k]] == -1) { owner[ci[k]] = -2; fifo[tail++] = ci[k]; }
            while (head < tail && size < block_size) {
                const index_t r = fifo[head++];
                if (owner[r] >= 0) continue;
                owner[r] = id;
                members.push_back(r);
                if (++size >= block_size) break;
                for (index_t k = rp[r]; k < rp[r + 1]; ++k)
                    if (owner[ci[k]] == -1) { owner[ci[k]] = -2; fifo[tail++] = ci[k]; }
            }
            for (; head < tail; ++head)
                if (owner[fifo[head]] == -2) owner[fifo[head]] = -1;
        }
        for (index_t i = 0; i < n; ++i) {
            if (owner[i] < 0) {
                if (first.empty()) first.push_back(0);
                owner[i] = index_t(first.size()) - 1;
                members.push_back(i);
            }
        }
        const index_t nb = index_t(first.size());
        first.push_back(index_t(members.size()));

        // ---- block graph (first-encounter neighbour order) ---------------
        std::vector<index_t> stamp(nb, -1);
        std::vector<index_t> grp(nb + 1, 0), gci;
        gci.reserve(size_t(nb) * 8);
        for (index_t b = 0; b < nb; ++b) {
            for (index_t q = first[b]; q < first[b + 1]; ++q) {
                const index_t r = members[q];
                for (index_t k = rp[r]; k < rp[r + 1]; ++k) {
                    const index_t o = owner[ci[k]];
                    if (o != b && stamp[o] != b) { stamp[o] = b; gci.push_back(o); }
                }
            }
            grp[b + 1] = index_t(gci.size());
        }

        // ---- colouring ----------------------------------------------------
        int ncol = target_colors;
        for (index_t b = 0; b < nb; ++b) {
            int lower = 0;
            for (index_t k = grp[b]; k < grp[b + 1]; ++k) lower += gci[k] < b;
            if (lower >= ncol) ncol = lower + 1;
        }
        std::vector<index_t> color(nb, -1), mark(size_t(ncol) + 1, -1);
        std::vector<index_t> per_color(size_t(ncol) + 1, 0);
        for (index_t b = 0; b < nb; ++b) {
            for (index_t k = grp[b]; k < grp[b + 1]; ++k)
                if (gci[k] < b) mark[color[gci[k]]] = b;
            int c = 0;
            while (c < ncol && mark[c] == b) ++c;
            if (c == ncol) {                      // every colour taken: open one
                ++ncol;
                mark.push_back(-1);
                per_color.push_back(0);
            }
            color[b] = c;
            ++per_color[c];
        }
        std::vector<index_t> remap(ncol, -1);
        index_t nc = 0;
        for (int c = 0; c < ncol; ++c) if (per_color[c] > 0) remap[c] = nc++;

        // ---- renumbering --------------------------------------------------
        color_offsets.assign(size_t(nc) + 1, 0);
        for (int c = 0; c < ncol; ++c) if (remap[c] >= 0) color_offsets[remap[c] + 1] = per_color[c];
        for (index_t c = 0; c < nc; ++c) color_offsets[c + 1] += color_offsets[c];
        std::vector<index_t> cursor(color_offsets.begin(), color_offsets.end() - 1);
        std::vector<index_t> slot(nb);
        for (index_t b = 0; b < nb; ++b) slot[cursor[remap[color[b]]]++] = b;
        block_offsets.assign(size_t(nb) + 1, 0);
        ordering.assign(n, 0);
        reverse_ordering.assign(n, 0);
        index_t next = 0;
        for (index_t g = 0; g < nb; ++g) {
            const index_t b = slot[g];
            index_t* seg = members.data() + first[b];
            const index_t len = first[b + 1] - first[b];
            std::sort(seg, seg + len);
            block_offsets[g] = next;
            for (index_t q = 0; q < len; ++q) {
                ordering[seg[q]] = next;
                reverse_ordering[next] = seg[q];
                ++next;
            }
        }
        block_offsets[nb] = next;
    }
};

}  // namespace sparsesolv_amd
