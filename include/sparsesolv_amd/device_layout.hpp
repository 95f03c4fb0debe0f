// device_layout.hpp -- HBM layouts of the operator and of the IC factor
// (host C++17, no device code).
//
// Everything lives in ABMC order: the Krylov vectors, the operator A and
// both triangular factors.  Two layouts are built here.
//
// 1. SpmvLayout: SELL-64.  Rows are taken 64 at a time (one wavefront), in
//    the matrix's original order (a map gives each lane its row);
//    slice s stores K_s = max row length entry slots, slot-major, so lane l
//    of slot k reads val[(off_s + k)*64 + l] -- every load of the wavefront
//    is one contiguous 512 B (f64) / 1 KiB (c128) segment.  Short rows are
//    padded with (value 0, column = own row).  Entries keep the CSR column
//    order of each row.
//
// 2. SweepLayout (one per direction).  The ABMC schedule is
//    colour -> blocks -> rows; blocks of one colour are independent, rows of
//    a block are sequential.  Each row of a regular block (at most BR rows)
//    gets a lane; the rows of a block sit in consecutive lanes of one slice,
//    within one run of kBlockAlign lanes, in processing order (ascending rows for the forward sweep over L,
//    descending for the backward sweep over L^T).  Blocks are sorted by
//    their longest row and packed into 64-lane slices; a slice holds
//    K = its longest row entry slots, laid out like SpmvLayout, and every
//    slice of a colour starts stride = the colour's largest K slots after
//    the previous one, so a kernel finds a slice's first entries from the
//    colour's first slot and stride alone (no offsets load before them);
//    each lane's info word carries its position in the block and its
//    slice's K.  A column c >= 0 is a finished row of an earlier phase
//    (gathered from HBM); c < 0 refers to position -1-c of the same block,
//    resolved by wavefront shuffles once that row is final.
//    The ABMC aggregation can leave one oversized last block (the rows left
//    unassigned once ceil(n/B) blocks exist).  Such a block ("big block")
//    is scheduled by its own dependency levels: rows grouped level by level,
//    entries as CSR, internal references encoded -1-p with p the
//    dependency's position in that level order.
#pragma once

#include <algorithm>
#include <cstdint>
#include <numeric>
#include <stdexcept>
#include <cstring>
#include <vector>

#include "abmc_ordering.hpp"
#include "ic_factor.hpp"

namespace sparsesolv_amd {

constexpr int kWave = 64;
// column of a padding slot in the sweep layouts: neither a row (>= 0) nor an
// in-block position (-1 - position, position < 64)
constexpr index_t kPadCol = -(1 << 30);
// steps of the straggler walk prefetched ahead, at most (the step lists are
// padded to a multiple of it, plus as many idle steps at the end; a kernel
// prefetching 8 or 16 steps fits both)
constexpr int kWalkDepth = 16;
// one 16-byte word of a packed walk record
struct WalkWord {
    unsigned char b[16];
};
// 16-byte words of a walk record for kmax internal entries of fs-byte
// values: kmax values, D^-1, then int32 row, position and kmax positions
constexpr int walk_record_words(int kmax, int fs) { return ((kmax + 1) * fs + (2 + kmax) * 4 + 15) / 16; }
// entries a sweep kernel loads per batch; the entry arrays end with this
// many padding slots so a batch never needs a bounds guard
constexpr int kSweepChunk = 16;
// the rows of a regular block (at most 16) never straddle a multiple of
// kBlockAlign lanes of a slice: a sweep that gives a wavefront 16 rows of a
// slice (several lanes a row) resolves every block within that wavefront
constexpr int kBlockAlign = 16;
// blocks a colour's longest-row sort reorders among: 16 blocks (64 rows of
// 4-row blocks) keep the slices that run together spatially close
// (64^3 curl-curl ICCG 72.0 -> 69.0 ms per 300 iterations, 48^3 Laplacian
// 11.5 -> 10.9 ms against a whole-colour sort; 16 right-hand sides
// unchanged within 1%)
constexpr size_t kSortWindow = 16;

template <typename F>
struct SpmvLayout {
    index_t n = 0, nslices = 0;
    std::vector<index_t> slice_off;   // nslices+1, in entry slots
    std::vector<F> val;               // slots*64
    std::vector<index_t> col;
    std::vector<index_t> row;         // nslices*64: the lane's row (-1: none)

    // Slices take the rows in the order `order` (order[k] = the row at
    // position k; empty: ascending).  With order = the matrix's original
    // numbering, a wavefront's rows, and the workgroups running at one time,
    // cover a compact part of the domain, so their gathers of x come from a
    // compact part of it too; in the colour order the rows in flight span
    // whole colours and reference nearly all of x.
    template <typename T>
    void build(const CsrHost<T>& A, const std::vector<index_t>& order = {})
    {
        n = A.n;
        nslices = (n + kWave - 1) / kWave;
        auto row_at = [&](index_t k) { return order.empty() ? k : order[size_t(k)]; };
        slice_off.assign(size_t(nslices) + 1, 0);
        row.assign(size_t(nslices) * kWave, -1);
        for (index_t s = 0; s < nslices; ++s) {
            index_t K = 0;
            for (index_t k = s * kWave; k < std::min<index_t>(n, (s + 1) * kWave); ++k) {
                const index_t r = row_at(k);
                row[size_t(k)] = r;
                K = std::max(K, A.rp[r + 1] - A.rp[r]);
            }
            slice_off[s + 1] = slice_off[s] + K;
        }
        const size_t slots = size_t(slice_off[nslices]) * kWave;
        val.assign(slots, F(0));
        col.assign(slots, 0);
        for (index_t s = 0; s < nslices; ++s) {
            const index_t K = slice_off[s + 1] - slice_off[s];
            for (int l = 0; l < kWave; ++l) {
                const index_t r = row[size_t(s) * kWave + l];
                for (index_t k = 0; k < K; ++k) {
                    const size_t at = (size_t(slice_off[s]) + k) * kWave + l;
                    if (r >= 0 && k < A.rp[r + 1] - A.rp[r]) {
                        val[at] = static_cast<F>(A.v[A.rp[r] + k]);
                        col[at] = A.ci[A.rp[r] + k];
                    } else {
                        col[at] = r >= 0 ? r : 0;   // padding: value 0
                    }
                }
            }
        }
    }
};

// The first row of every system in combine_spmv's layout, and the end.
template <typename F>
std::vector<index_t> combined_rows(const std::vector<const SpmvLayout<F>*>& parts, index_t align)
{
    std::vector<index_t> row0(1, 0);
    for (const SpmvLayout<F>* P : parts)
        row0.push_back(row0.back() + (P->nslices + align - 1) / align * align * kWave);
    return row0;
}

// Several systems in one SELL-64 layout: system s occupies the rows
// [row0[s], row0[s+1]); each system's slice count is padded to a multiple of
// `align` slices (empty slices) so no group of `align` slices spans two
// systems.  Columns are shifted by the system's first row.  The sizes are
// summed first, then the systems are copied side by side.
template <typename F>
SpmvLayout<F> combine_spmv(const std::vector<const SpmvLayout<F>*>& parts, index_t align,
                           std::vector<index_t>& row0)
{
    SpmvLayout<F> out;
    const size_t np = parts.size();
    std::vector<index_t> slice0(np + 1, 0);     // first combined slice of each system
    std::vector<size_t> slot0(np + 1, 0);       // its first entry slot
    row0 = combined_rows(parts, align);
    for (size_t i = 0; i < np; ++i) {
        const SpmvLayout<F>& P = *parts[i];
        slice0[i + 1] = slice0[i] + (P.nslices + align - 1) / align * align;
        slot0[i + 1] = slot0[i] + size_t(P.slice_off[P.nslices]);
    }
    out.nslices = slice0[np];
    out.n = row0.back();
    out.slice_off.assign(size_t(out.nslices) + 1, 0);
    out.row.assign(std::max<size_t>(size_t(out.nslices) * kWave, kWave), -1);
    out.val.resize(std::max<size_t>(slot0[np] * kWave, kWave));
    out.col.resize(out.val.size());
    if (slot0[np] == 0) { out.val.assign(kWave, F(0)); out.col.assign(kWave, 0); }
    parallel_items(index_t(np), [&](index_t i) {
        const SpmvLayout<F>& P = *parts[size_t(i)];
        const index_t r0 = row0[size_t(i)];
        const size_t base = slot0[size_t(i)] * kWave;
        const size_t len = size_t(P.slice_off[P.nslices]) * kWave;
        std::copy(P.val.begin(), P.val.begin() + len, out.val.begin() + base);
        for (size_t q = 0; q < len; ++q) out.col[base + q] = P.col[q] + r0;
        index_t* so = out.slice_off.data() + slice0[size_t(i)];
        for (index_t s = slice0[size_t(i)]; s < slice0[size_t(i) + 1]; ++s, ++so)
            so[1] = index_t(slot0[size_t(i)]) + P.slice_off[std::min(s - slice0[size_t(i)] + 1, P.nslices)];
        index_t* rw = out.row.data() + size_t(slice0[size_t(i)]) * kWave;
        for (size_t q = 0; q < size_t(P.nslices) * kWave; ++q) rw[q] = P.row[q] >= 0 ? P.row[q] + r0 : -1;
    });
    return out;
}

// One triangular sweep (forward over L or backward over L^T).
template <typename F>
struct SweepLayout {
    int BR = 1;                         // rows per lane (power of two)
    index_t nphases = 0;                // colours, in processing order
    std::vector<index_t> phase_slices;  // nphases+1: slice ranges
    std::vector<index_t> lane_row;      // nslices*64: row of the lane (-1: idle)
    std::vector<index_t> lane_info;     // nslices*64: position in the block | slice's K << 8
    std::vector<index_t> slice_slot;    // nslices+1: first entry slot of each slice, end
    std::vector<index_t> phase_slot;    // nphases: first entry slot of the colour's slices
    std::vector<index_t> phase_stride;  // nphases: entry slots from one slice to the next
    std::vector<F> val;                 // slots*64
    std::vector<index_t> col;
    // per colour, rows of the regular slices, their factor entries counting
    // the diagonal, and the distinct rows of earlier colours their entries
    // read (the algorithmic size of one colour kernel: every operand once)
    std::vector<index_t> phase_rows;
    std::vector<int64_t> phase_entries;
    std::vector<int64_t> phase_ext_rows;

    // big blocks (level scheduled), each owned by one workgroup
    std::vector<index_t> phase_big;     // nphases+1: big-block ranges
    std::vector<index_t> big_pos;       // nbig+1: offsets into big_rows
    std::vector<index_t> big_lvl;       // level offsets, concatenated
    std::vector<index_t> big_lvl_off;   // nbig+1: offsets into big_lvl
    std::vector<index_t> big_rows;      // rows in level order
    std::vector<index_t> big_rp;        // per big row (global position), CSR
    std::vector<index_t> big_ci;
    std::vector<F> big_v;
    index_t max_big_rows = 0;
    index_t max_levels = 0;

    // Single-wavefront walk of the big blocks.  Each level is cut into steps
    // of at most 64 rows; slot t = step*64 + lane.  Per slot: the row, its
    // position in level order (index of its value in LDS), D^-1, the
    // external entries (CSR over slots, summed by the whole workgroup
    // first) and exactly walk_kmax internal entries (value, position;
    // padding points at the always-zero position max_big_rows).  Idle slots
    // have position -1; the kernel sends their value to a scratch LDS slot.
    int walk_kmax = 0;                   // 0: walk unavailable (fallback)
    std::vector<index_t> walk_step_off;  // nbig+1
    std::vector<index_t> walk_row;       // slots (-1 idle)
    std::vector<index_t> walk_pos;       // slots (-1 idle)
    std::vector<index_t> walk_ext_rp;    // slots+1
    std::vector<index_t> walk_ext_ci;
    std::vector<F> walk_ext_v;
    std::vector<F> walk_invd;            // slots
    std::vector<F> walk_v;               // (step*kmax + k)*64 + lane
    std::vector<index_t> walk_ix;
    // The same per-slot operands packed for the device (pack_walk): one
    // record of walk_rec_words 16-byte words per slot -- the kmax values,
    // D^-1, then int32 row, position and the kmax LDS positions -- stored
    // word-major per step (word w of lane l of step s at (s*words + w)*64 +
    // l), so each word of a step is one coalesced load and a step costs
    // walk_rec_words loads instead of 2*kmax + 3.
    std::vector<WalkWord> walk_rec;
    int walk_rec_words = 0;

    index_t nslices() const { return phase_slices.empty() ? 0 : phase_slices.back(); }
    // lane info word: the row's position in its block (< 64) and the
    // slice's entry slots K
    static constexpr int kInfoShift = 8;
    static index_t info_word(index_t pos, index_t K) { return pos | (K << kInfoShift); }
    static index_t info_pos(index_t w) { return w & ((1 << kInfoShift) - 1); }
    static index_t info_len(index_t w) { return w >> kInfoShift; }

    // forward: rows of L below the diagonal; backward: rows of U = L^T above.
    template <typename T>
    void build(const AbmcSchedule& S, const CsrHost<T>& M, bool forward, int block_size,
               const std::vector<F>& invd)
    {
        const index_t nc = S.num_colors();
        const index_t nb = S.num_blocks();
        index_t maxreg = 1;
        // only the last block can exceed block_size (aggregation stragglers)
        for (index_t b = 0; b < nb; ++b) {
            const index_t sz = S.block_offsets[b + 1] - S.block_offsets[b];
            if (sz <= block_size) maxreg = std::max(maxreg, sz);
        }
        BR = 1;
        while (BR < maxreg) BR *= 2;
        if (BR > 16) BR = 0;            // wide blocks: everything level scheduled
        nphases = nc;
        phase_slices.assign(size_t(nc) + 1, 0);
        phase_big.assign(size_t(nc) + 1, 0);
        big_pos.assign(1, 0);
        big_lvl_off.assign(1, 0);
        big_lvl.clear(); big_rows.clear(); big_rp.assign(1, 0); big_ci.clear(); big_v.clear();
        lane_row.clear(); lane_info.clear(); slice_slot.assign(1, 0);
        phase_slot.assign(size_t(nc), 0);
        phase_stride.assign(size_t(nc), 0);
        val.clear(); col.clear();
        max_big_rows = 0; max_levels = 0;

        // strictly-off-diagonal entries of row i in sweep direction
        auto row_span = [&](index_t i, index_t& a, index_t& e) {
            if (forward) { a = M.rp[i]; e = M.rp[i + 1] - 1; }      // diag last
            else { a = M.rp[i] + 1; e = M.rp[i + 1]; }               // diag first
        };
        for (index_t p = 0; p < nc; ++p) {
            const index_t c = forward ? p : nc - 1 - p;
            std::vector<index_t> regular, big;
            for (index_t b = S.color_offsets[c]; b < S.color_offsets[c + 1]; ++b) {
                const index_t sz = S.block_offsets[b + 1] - S.block_offsets[b];
                (BR > 0 && sz <= BR ? regular : big).push_back(b);
            }
            // Row lanes: the rows of a block occupy consecutive lanes of one
            // slice, in processing order (position 0 first).  Blocks are sorted
            // by their longest row so that a slice's slot count (its longest
            // row) wastes little padding.
            auto longest = [&](index_t b) -> index_t {
                index_t m = 0;
                for (index_t i = S.block_offsets[b]; i < S.block_offsets[b + 1]; ++i) {
                    index_t a, e;
                    row_span(i, a, e);
                    m = std::max(m, e - a);
                }
                return m;
            };
            std::vector<std::pair<index_t, index_t>> keyed;
            keyed.reserve(regular.size());
            for (index_t b : regular) keyed.emplace_back(longest(b), b);
            // sorted within windows of kSortWindow consecutive blocks: the
            // slices that run together keep the blocks' (breadth-first,
            // hence spatial) order, so their gathers stay close
            for (size_t w0 = 0; w0 < keyed.size(); w0 += kSortWindow)
                std::stable_sort(keyed.begin() + w0,
                                 keyed.begin() + std::min(keyed.size(), w0 + kSortWindow),
                                 [](const auto& x, const auto& y) { return x.first > y.first; });
            // fill slices with whole blocks, then lay them out stride slots
            // apart
            struct Slice {
                index_t row[kWave], pos[kWave], b0[kWave], b1[kWave];
                index_t K = 0;
            };
            std::vector<Slice> slices;
            size_t q = 0;
            while (q < keyed.size()) {
                Slice sl;
                std::fill(sl.row, sl.row + kWave, -1);
                std::fill(sl.pos, sl.pos + kWave, 0);
                std::fill(sl.b0, sl.b0 + kWave, 0);
                std::fill(sl.b1, sl.b1 + kWave, 0);
                int used = 0;
                while (q < keyed.size()) {
                    const index_t b = keyed[q].second;
                    const index_t b0 = S.block_offsets[b], b1 = S.block_offsets[b + 1];
                    const int sz = int(b1 - b0);
                    // a block never crosses a multiple of kBlockAlign lanes
                    if (used % kBlockAlign + sz > kBlockAlign) used += kBlockAlign - used % kBlockAlign;
                    if (used + sz > kWave) break;
                    for (int r = 0; r < sz; ++r) {
                        sl.row[used + r] = forward ? b0 + r : b1 - 1 - r;
                        sl.pos[used + r] = r;
                        sl.b0[used + r] = b0;
                        sl.b1[used + r] = b1;
                    }
                    used += sz;
                    ++q;
                }
                for (int l = 0; l < kWave; ++l) {
                    if (sl.row[l] < 0) continue;
                    index_t a, e;
                    row_span(sl.row[l], a, e);
                    sl.K = std::max(sl.K, e - a);
                }
                slices.push_back(sl);
            }
            const index_t nsl = index_t(slices.size());
            index_t stride = 0;
            for (const Slice& sl : slices) stride = std::max(stride, sl.K);
            const index_t first = slice_slot.back();
            phase_slot[p] = first;
            phase_stride[p] = stride;
            slice_slot.pop_back();
            for (index_t k = 0; k <= nsl; ++k) slice_slot.push_back(first + k * stride);
            val.resize(size_t(first + nsl * stride) * kWave, F(0));
            col.resize(size_t(first + nsl * stride) * kWave, kPadCol);
            for (index_t k = 0; k < nsl; ++k) {
                const Slice& sl = slices[size_t(k)];
                const size_t base = size_t(first + k * stride) * kWave;
                for (int l = 0; l < kWave; ++l) {
                    lane_row.push_back(sl.row[l]);
                    lane_info.push_back(info_word(sl.pos[l], sl.K));
                    if (sl.row[l] < 0) continue;
                    index_t a, e;
                    row_span(sl.row[l], a, e);
                    const index_t b0 = sl.b0[l], b1 = sl.b1[l];
                    for (index_t k2 = a; k2 < e; ++k2) {
                        const index_t cc = M.ci[k2];
                        const bool internal = cc >= b0 && cc < b1;
                        const index_t pos = forward ? cc - b0 : b1 - 1 - cc;
                        const size_t at = base + size_t(k2 - a) * kWave + l;
                        val[at] = static_cast<F>(M.v[k2]);
                        col[at] = internal ? -1 - pos : cc;
                    }
                }
            }
            phase_slices[p + 1] = phase_slices[p] + nsl;

            // level-scheduled blocks
            for (index_t b : big) {
                const index_t b0 = S.block_offsets[b], b1 = S.block_offsets[b + 1];
                const index_t sz = b1 - b0;
                std::vector<index_t> level(sz, 0);
                index_t nlev = 0;
                for (index_t q = 0; q < sz; ++q) {
                    const index_t i = forward ? b0 + q : b1 - 1 - q;
                    index_t a, e;
                    row_span(i, a, e);
                    index_t lv = 0;
                    for (index_t k = a; k < e; ++k) {
                        const index_t cc = M.ci[k];
                        if (cc >= b0 && cc < b1) {
                            const index_t dq = forward ? cc - b0 : b1 - 1 - cc;
                            lv = std::max(lv, level[dq] + 1);
                        }
                    }
                    level[q] = lv;
                    nlev = std::max(nlev, lv + 1);
                }
                // counting sort of positions by level (stable in q)
                std::vector<index_t> loff(size_t(nlev) + 1, 0);
                for (index_t q = 0; q < sz; ++q) loff[level[q] + 1]++;
                for (index_t v = 0; v < nlev; ++v) loff[v + 1] += loff[v];
                std::vector<index_t> at(loff.begin(), loff.end() - 1), posof(sz);
                std::vector<index_t> order(sz);
                for (index_t q = 0; q < sz; ++q) { posof[q] = at[level[q]]++; order[posof[q]] = q; }
                for (index_t v = 0; v <= nlev; ++v) big_lvl.push_back(loff[v]);
                big_lvl_off.push_back(index_t(big_lvl.size()));
                for (index_t p2 = 0; p2 < sz; ++p2) {
                    const index_t q = order[p2];
                    const index_t i = forward ? b0 + q : b1 - 1 - q;
                    big_rows.push_back(i);
                    index_t a, e;
                    row_span(i, a, e);
                    for (index_t k = a; k < e; ++k) {
                        const index_t cc = M.ci[k];
                        const bool internal = cc >= b0 && cc < b1;
                        const index_t dq = forward ? cc - b0 : b1 - 1 - cc;
                        big_ci.push_back(internal ? -1 - posof[dq] : cc);
                        big_v.push_back(static_cast<F>(M.v[k]));
                    }
                    big_rp.push_back(index_t(big_ci.size()));
                }
                big_pos.push_back(index_t(big_rows.size()));
                max_big_rows = std::max(max_big_rows, sz);
                max_levels = std::max(max_levels, nlev);
            }
            phase_big[p + 1] = phase_big[p] + index_t(big.size());
        }
        phase_rows.assign(size_t(nphases), 0);
        phase_entries.assign(size_t(nphases), 0);
        phase_ext_rows.assign(size_t(nphases), 0);
        std::vector<index_t> seen(size_t(M.n), -1);
        for (index_t p = 0; p < nphases; ++p)
            for (index_t sl = phase_slices[p]; sl < phase_slices[p + 1]; ++sl) {
                for (int lane = 0; lane < kWave; ++lane) {
                    const index_t row = lane_row[size_t(sl) * kWave + lane];
                    if (row < 0) continue;
                    ++phase_rows[p];
                    phase_entries[p] += M.rp[row + 1] - M.rp[row];   // diagonal included
                }
                // external entries (columns >= 0 among the slice's K slots)
                const index_t K = info_len(lane_info[size_t(sl) * kWave]);
                const size_t at = size_t(slice_slot[sl]) * kWave;
                for (size_t q = 0; q < size_t(K) * kWave; ++q) {
                    const index_t c = col[at + q];
                    if (c >= 0 && seen[size_t(c)] != p) {
                        seen[size_t(c)] = p;
                        ++phase_ext_rows[p];
                    }
                }
            }
        build_walk(M, forward, invd);
        val.resize(val.size() + size_t(kSweepChunk) * kWave, F(0));
        col.resize(col.size() + size_t(kSweepChunk) * kWave, kPadCol);
        if (big_ci.empty()) { big_ci.push_back(0); big_v.push_back(F(0)); }
        if (big_rows.empty()) big_rows.push_back(0);
        if (big_lvl.empty()) big_lvl.push_back(0);
        if (lane_row.empty()) { lane_row.assign(kWave, -1); lane_info.assign(kWave, 0); }
    }

private:
    void push_idle_step()
    {
        const int K = walk_kmax;
        walk_v.resize(walk_v.size() + size_t(K) * kWave, F(0));
        walk_ix.resize(walk_ix.size() + size_t(K) * kWave, max_big_rows);
        for (int lane = 0; lane < kWave; ++lane) {
            walk_row.push_back(-1);
            walk_pos.push_back(-1);
            walk_invd.push_back(F(0));
            walk_ext_rp.push_back(index_t(walk_ext_ci.size()));
        }
    }

    template <typename T>
    void build_walk(const CsrHost<T>& M, bool forward, const std::vector<F>& invd)
    {
        const index_t nbig = index_t(big_pos.size()) - 1;
        walk_step_off.assign(1, 0);
        walk_row.clear(); walk_pos.clear(); walk_ext_rp.assign(1, 0); walk_ext_ci.clear();
        walk_ext_v.clear(); walk_invd.clear(); walk_v.clear(); walk_ix.clear();
        // internal entries per row: the ones encoded as positions (< 0)
        int kmax = 0;
        for (index_t gp = 0; gp + 1 < index_t(big_rp.size()); ++gp) {
            int k = 0;
            for (index_t q = big_rp[gp]; q < big_rp[gp + 1]; ++q) k += big_ci[q] < 0;
            kmax = std::max(kmax, k);
        }
        walk_kmax = nbig == 0 ? 0 : (kmax <= 2 ? 2 : kmax <= 4 ? 4 : kmax <= 8 ? 8 : 0);
        if (walk_kmax == 0) {
            walk_step_off.assign(size_t(nbig) + 1, 0);
        } else {
            const int K = walk_kmax;
            for (index_t bb = 0; bb < nbig; ++bb) {
                const index_t p0 = big_pos[bb];
                const index_t* lv = big_lvl.data() + big_lvl_off[bb];
                const index_t nlev = big_lvl_off[bb + 1] - big_lvl_off[bb] - 1;
                index_t nsteps = 0;
                for (index_t l = 0; l < nlev; ++l) {
                    for (index_t q0 = lv[l]; q0 < lv[l + 1]; q0 += kWave) {
                        const index_t q1 = std::min<index_t>(lv[l + 1], q0 + kWave);
                        const size_t vbase = walk_v.size();
                        walk_v.resize(vbase + size_t(K) * kWave, F(0));
                        walk_ix.resize(vbase + size_t(K) * kWave, max_big_rows);
                        for (int lane = 0; lane < kWave; ++lane) {
                            const index_t q = q0 + lane;
                            if (q >= q1) {
                                walk_row.push_back(-1); walk_pos.push_back(-1);
                                walk_invd.push_back(F(0));
                                walk_ext_rp.push_back(index_t(walk_ext_ci.size()));
                                continue;
                            }
                            const index_t gp = p0 + q;
                            const index_t row = big_rows[gp];
                            walk_row.push_back(row);
                            walk_pos.push_back(q);
                            walk_invd.push_back(invd[row]);
                            int k = 0;
                            for (index_t e = big_rp[gp]; e < big_rp[gp + 1]; ++e) {
                                if (big_ci[e] >= 0) {
                                    walk_ext_ci.push_back(big_ci[e]);
                                    walk_ext_v.push_back(big_v[e]);
                                } else {
                                    const size_t at = vbase + size_t(k) * kWave + lane;
                                    walk_v[at] = big_v[e];
                                    walk_ix[at] = -1 - big_ci[e];
                                    ++k;
                                }
                            }
                            walk_ext_rp.push_back(index_t(walk_ext_ci.size()));
                        }
                        ++nsteps;
                    }
                }
                // idle steps up to a multiple of kWalkDepth: the walk kernel's
                // loop then has no tail and no branch around its prefetches
                while (nsteps % kWalkDepth) {
                    push_idle_step();
                    ++nsteps;
                }
                walk_step_off.push_back(walk_step_off.back() + nsteps);
            }
            // the prefetch of the last kWalkDepth steps reads this idle tail
            for (int k = 0; k < kWalkDepth; ++k) push_idle_step();
        }
        if (walk_row.empty()) { walk_row.assign(1, -1); walk_pos.assign(1, -1); walk_invd.assign(1, F(0)); walk_ext_rp.push_back(0); }
        if (walk_ext_ci.empty()) { walk_ext_ci.push_back(0); walk_ext_v.push_back(F(0)); }
        if (walk_v.empty()) { walk_v.push_back(F(0)); walk_ix.push_back(0); }
        pack_walk();
        (void)M; (void)forward;
    }

    void pack_walk()
    {
        const int K = walk_kmax;
        walk_rec_words = K > 0 ? walk_record_words(K, int(sizeof(F))) : 0;
        if (K == 0) {
            walk_rec.assign(1, WalkWord{});
            return;
        }
        const size_t W = size_t(walk_rec_words), nsteps = walk_row.size() / kWave;
        const size_t FS = sizeof(F);
        walk_rec.assign(nsteps * W * kWave, WalkWord{});
        std::vector<unsigned char> rec(W * 16);
        for (size_t st = 0; st < nsteps; ++st)
            for (int l = 0; l < kWave; ++l) {
                std::fill(rec.begin(), rec.end(), 0);
                const size_t t = st * kWave + size_t(l);
                for (int k = 0; k < K; ++k)
                    std::memcpy(&rec[size_t(k) * FS], &walk_v[(st * size_t(K) + size_t(k)) * kWave + size_t(l)], FS);
                std::memcpy(&rec[size_t(K) * FS], &walk_invd[t], FS);
                const size_t ib = size_t(K + 1) * FS;
                std::memcpy(&rec[ib], &walk_row[t], 4);
                std::memcpy(&rec[ib + 4], &walk_pos[t], 4);
                for (int k = 0; k < K; ++k)
                    std::memcpy(&rec[ib + 8 + size_t(k) * 4], &walk_ix[(st * size_t(K) + size_t(k)) * kWave + size_t(l)], 4);
                for (size_t w = 0; w < W; ++w)
                    std::memcpy(walk_rec[(st * W + w) * kWave + size_t(l)].b, &rec[w * 16], 16);
            }
    }

public:
    // Several systems' sweeps in one layout: phase p holds phase p of every
    // system that has one (systems are independent, so their colours can
    // run side by side); rows and external columns are shifted by the
    // system's first row (row0[s]).  slice_sys / walk_sys name the system of
    // every slice / oversized block.  Oversized blocks must use the walk
    // (walk_kmax > 0); the level-kernel fallback arrays are left empty.
    static SweepLayout combine(const std::vector<const SweepLayout*>& parts,
                               const std::vector<index_t>& row0, std::vector<int>& slice_sys,
                               std::vector<int>& walk_sys)
    {
        SweepLayout out;
        slice_sys.clear();
        walk_sys.clear();
        bool any_big = false;
        for (const SweepLayout* P : parts) {
            if (P->BR == 0) throw std::runtime_error("batch: blocks wider than 16 rows are not supported");
            out.BR = std::max(out.BR, P->BR);
            out.nphases = std::max(out.nphases, P->nphases);
            out.max_big_rows = std::max(out.max_big_rows, P->max_big_rows);
            out.max_levels = std::max(out.max_levels, P->max_levels);
            if (!P->phase_big.empty() && P->phase_big.back() > 0) {
                if (P->walk_kmax == 0)
                    throw std::runtime_error("batch: an oversized block needs more than 8 in-block entries per row");
                any_big = true;
                out.walk_kmax = std::max(out.walk_kmax, P->walk_kmax);
            }
        }
        const int K = out.walk_kmax;
        const size_t np = parts.size();
        const index_t nph = out.nphases;
        out.phase_slices.assign(1, 0);
        out.phase_big.assign(1, 0);
        out.slice_slot.assign(1, 0);
        out.phase_slot.assign(size_t(nph), 0);
        out.phase_stride.assign(size_t(nph), 0);
        out.walk_step_off.assign(1, 0);
        out.phase_rows.assign(size_t(nph), 0);
        out.phase_entries.assign(size_t(nph), 0);
        out.phase_ext_rows.assign(size_t(nph), 0);
        // Sizes first: phase p holds, system by system, the slices of every
        // system's phase p (one stride per phase: its longest slice), then
        // the walk steps of their oversized blocks.  Per (phase, system):
        // its first combined slice, walk step and external walk entry.
        std::vector<index_t> first_slice(size_t(nph) * np, 0), first_step(size_t(nph) * np, 0);
        std::vector<int64_t> first_ext(size_t(nph) * np, 0);
        index_t nsl_all = 0, steps_all = 0;
        int64_t ext_all = 0;
        for (index_t p = 0; p < nph; ++p) {
            index_t stride = 0, nsl = 0;
            for (size_t si = 0; si < np; ++si) {
                const SweepLayout& P = *parts[si];
                first_slice[size_t(p) * np + si] = nsl_all + nsl;
                if (p >= P.nphases) continue;
                for (index_t sl = P.phase_slices[p]; sl < P.phase_slices[p + 1]; ++sl, ++nsl) {
                    stride = std::max(stride, info_len(P.lane_info[size_t(sl) * kWave]));
                    slice_sys.push_back(int(si));
                }
                out.phase_rows[p] += P.phase_rows[p];
                out.phase_entries[p] += P.phase_entries[p];
                out.phase_ext_rows[p] += P.phase_ext_rows[p];
            }
            const index_t first = out.slice_slot.back();
            out.phase_slot[p] = first;
            out.phase_stride[p] = stride;
            out.slice_slot.pop_back();
            for (index_t k = 0; k <= nsl; ++k) out.slice_slot.push_back(first + k * stride);
            nsl_all += nsl;
            out.phase_slices.push_back(nsl_all);
            for (size_t si = 0; si < np; ++si) {
                const SweepLayout& P = *parts[si];
                first_step[size_t(p) * np + si] = steps_all;
                first_ext[size_t(p) * np + si] = ext_all;
                if (p >= P.nphases) continue;
                for (index_t bb = P.phase_big[p]; bb < P.phase_big[p + 1]; ++bb) {
                    const index_t st0 = P.walk_step_off[bb], st1 = P.walk_step_off[bb + 1];
                    steps_all += st1 - st0;
                    ext_all += P.walk_ext_rp[size_t(st1) * kWave] - P.walk_ext_rp[size_t(st0) * kWave];
                    out.walk_step_off.push_back(out.walk_step_off.back() + (st1 - st0));
                    walk_sys.push_back(int(si));
                }
            }
            out.phase_big.push_back(index_t(walk_sys.size()));
        }
        const size_t slots = size_t(out.slice_slot.back());
        out.lane_row.assign(size_t(nsl_all) * kWave, -1);
        out.lane_info.assign(size_t(nsl_all) * kWave, 0);
        out.val.assign((slots + kSweepChunk) * kWave, F(0));
        out.col.assign((slots + kSweepChunk) * kWave, kPadCol);
        const size_t wslots = size_t(steps_all) * kWave;
        out.walk_row.assign(wslots, -1);
        out.walk_pos.assign(wslots, -1);
        out.walk_invd.assign(wslots, F(0));
        out.walk_ext_rp.assign(wslots + 1, 0);
        out.walk_ext_ci.assign(size_t(ext_all), 0);
        out.walk_ext_v.assign(size_t(ext_all), F(0));
        out.walk_v.assign(wslots * size_t(K), F(0));
        out.walk_ix.assign(wslots * size_t(K), out.max_big_rows);
        // then every system's part, the systems side by side
        parallel_items(index_t(np), [&](index_t sidx) {
            const size_t si = size_t(sidx);
            const SweepLayout& P = *parts[si];
            const index_t r0 = row0[si];
            for (index_t p = 0; p < std::min(nph, P.nphases); ++p) {
                index_t k = first_slice[size_t(p) * np + si];
                for (index_t sl = P.phase_slices[p]; sl < P.phase_slices[p + 1]; ++sl, ++k) {
                    const index_t Ks = info_len(P.lane_info[size_t(sl) * kWave]);
                    const size_t src = size_t(P.slice_slot[sl]) * kWave;
                    const size_t dst = size_t(out.slice_slot[k]) * kWave;
                    for (size_t q = 0; q < size_t(Ks) * kWave; ++q) {
                        const index_t c = P.col[src + q];
                        out.col[dst + q] = c >= 0 ? c + r0 : c;
                        out.val[dst + q] = P.val[src + q];
                    }
                    for (int l = 0; l < kWave; ++l) {
                        const size_t a = size_t(sl) * kWave + l, b = size_t(k) * kWave + l;
                        const index_t r = P.lane_row[a];
                        out.lane_row[b] = r >= 0 ? r + r0 : r;
                        out.lane_info[b] = P.lane_info[a];
                    }
                }
                if (P.phase_big[p + 1] == P.phase_big[p]) continue;
                const index_t st0 = P.walk_step_off[P.phase_big[p]];
                const index_t st1 = P.walk_step_off[P.phase_big[p + 1]];
                const size_t t0 = size_t(st0) * kWave;
                const size_t g0 = size_t(first_step[size_t(p) * np + si]) * kWave;
                const int64_t e0 = P.walk_ext_rp[t0];
                const int64_t ge = first_ext[size_t(p) * np + si];
                for (size_t t = t0; t < size_t(st1) * kWave; ++t) {
                    const size_t g = g0 + (t - t0);
                    const index_t r = P.walk_row[t];
                    out.walk_row[g] = r >= 0 ? r + r0 : r;
                    out.walk_pos[g] = P.walk_pos[t];
                    out.walk_invd[g] = P.walk_invd[t];
                    for (index_t e = P.walk_ext_rp[t]; e < P.walk_ext_rp[t + 1]; ++e) {
                        out.walk_ext_ci[size_t(ge + (e - e0))] = P.walk_ext_ci[e] + r0;
                        out.walk_ext_v[size_t(ge + (e - e0))] = P.walk_ext_v[e];
                    }
                    out.walk_ext_rp[g + 1] = index_t(ge + (P.walk_ext_rp[t + 1] - e0));
                }
                for (index_t st = st0; st < st1; ++st) {
                    const size_t vbase = (size_t(first_step[size_t(p) * np + si] + (st - st0))) * size_t(K) * kWave;
                    for (int k = 0; k < P.walk_kmax; ++k)
                        for (int l = 0; l < kWave; ++l) {
                            const size_t a = (size_t(st) * P.walk_kmax + k) * kWave + l;
                            const index_t ix = P.walk_ix[a];
                            out.walk_v[vbase + size_t(k) * kWave + l] = P.walk_v[a];
                            // the system's zero slot becomes the combined one
                            out.walk_ix[vbase + size_t(k) * kWave + l] =
                                ix == P.max_big_rows ? out.max_big_rows : ix;
                        }
                }
            }
        });
        if (any_big)
            for (int k = 0; k < kWalkDepth; ++k) out.push_idle_step();
        out.big_pos.assign(size_t(walk_sys.size()) + 1, 0);
        out.big_lvl_off.assign(size_t(walk_sys.size()) + 1, 0);
        out.big_lvl.assign(1, 0);
        out.big_rows.assign(1, 0);
        out.big_rp.assign(1, 0);
        out.big_ci.assign(1, 0);
        out.big_v.assign(1, F(0));
        if (out.lane_row.empty()) { out.lane_row.assign(kWave, -1); out.lane_info.assign(kWave, 0); }
        if (out.walk_row.empty()) { out.walk_row.assign(1, -1); out.walk_pos.assign(1, -1); out.walk_invd.assign(1, F(0)); out.walk_ext_rp.push_back(0); }
        if (out.walk_ext_ci.empty()) { out.walk_ext_ci.push_back(0); out.walk_ext_v.push_back(F(0)); }
        if (out.walk_v.empty()) { out.walk_v.push_back(F(0)); out.walk_ix.push_back(0); }
        out.pack_walk();
        return out;
    }
};

}  // namespace sparsesolv_amd
