// ic_impl.hpp -- the handle behind the C ABI: upload, IC apply, SpMV and the
// ICCG / COCR drivers on one MI355X.  Each ic_*.hip file instantiates it for
// one precision pair.
//
// Design (one GPU, one HIP stream per context):
//  * the host (host_setup.cpp) computes the ABMC schedule and the IC(0)
//    factor; this file uploads A (ABMC order, SELL-64), the forward and
//    backward sweep layouts and D^-1 once per ssa_ic_create;
//  * the Krylov loop runs entirely in ABMC order (b is gathered in, x is
//    gathered out), so the preconditioner applies no permutation;
//  * scalars never leave the device: control kernels reduce the
//    per-workgroup partials in index order and update SolverState; the host
//    only polls the stop status every few iterations;
//  * one iteration is captured into a hipGraph and replayed.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "host_plan.hpp"
#include "block_kernels.hpp"
#include "kernels.hpp"
#include "mrtr_kernels.hpp"
#include "ic_handle.hpp"
#include "sparsesolv_amd.h"

using ssa::index_t;

namespace ssa_impl {

// owning device buffer
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    // true when new memory was allocated
    bool alloc(size_t b)
    {
        if (b == bytes && p) return false;
        release();
        if (b == 0) b = 16;
        HIP_CHECK(hipMalloc(&p, b));
        bytes = b;
        return true;
    }
    // at least b bytes (kept when large enough); true when new memory was
    // allocated
    bool reserve(size_t b)
    {
        if (p && bytes >= b) return false;
        return alloc(b);
    }
    template <typename V>
    void upload(const std::vector<V>& v, hipStream_t s)
    {
        alloc(v.size() * sizeof(V));
        if (!v.empty()) HIP_CHECK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice, s));
    }
    template <typename V> V* as() const { return static_cast<V*>(p); }
};

constexpr int kMaxGrid = 1024;  // vector kernels: partial count bound
// sweep entries per batch with W right-hand sides per lane and `budget`
// bytes of gathered values per lane in flight (1 to 8)
template <typename P, int W, int budget> constexpr int chunk_of()
{
    constexpr int c = int(budget / (W * sizeof(P)));
    return c > 8 ? 8 : c < 1 ? 1 : c;
}
// sweep entries per batch: 16 for real, 8 for complex values; with W
// right-hand sides per lane at most 64 / W gathers and 512 bytes of
// gathered values per lane in flight -- 256 bytes at W = 16 with 8-byte
// values (the lane-a-row colour kernel, which blocks of right-hand sides
// now run only for an external part alone; when it ran their colours, a
// batch of four 16-value complex<float> rows took 202 VGPRs, 2 waves per
// SIMD, and two 126-143, 3-4 waves: eddy78 colour kernel 32.5 -> 31.9 us;
// complex<double> with one entry a batch: FP64 block 3.23 -> 3.29 s).  The
// layouts pad their entry arrays by kSweepChunk.
template <typename P, int W> constexpr int sweep_chunk()
{
    if (W == 1) return sizeof(P) > 8 ? 8 : 16;
    constexpr int c = chunk_of<P, W, (W >= 16 && sizeof(P) == 8) ? 256 : 512>();
    return c < 64 / W ? c : 64 / W;
}
// A colour of at most kSmallColourSlices slices (256 CUs x 4 SIMDs) leaves
// SIMDs idle: with several right-hand sides its wavefronts each get a
// workgroup (and a CU's address unit) of their own.  Eddy78, 16 right-hand
// sides: the forward colours of 210 and 52 slices 21 -> 14 and 19 -> 13 us
// with a lane a row, 9.4 and 8.4 us with four lanes a row.
constexpr index_t kSmallColourSlices = 1024;
// entries a batch for a lone wavefront per SIMD (the small-colour tail at 256
// threads): up to 1024 bytes of gathered values a lane (256 VGPRs and the
// AGPRs; 512 bytes for complex<double>, whose 8-row blocks spilled to
// scratch at 1024)
template <typename P, int W> constexpr int sweep_chunk_small()
{
    return W == 1 ? sweep_chunk<P, 1>() : chunk_of<P, W, (sizeof(P) > 8 ? 512 : 1024)>();
}
// lanes a row in the split colour kernel: as many as the row has 16-byte
// words, at most 4 (16 complex<float> values: 4 lanes of 32 bytes, one
// 128-B line a row per load pair)
template <typename P, int W> constexpr int split_lanes()
{
    constexpr int words = int(W * sizeof(P) / 16);
    return words >= 4 ? 4 : words >= 2 ? 2 : 1;
}
// entries a batch in the split colour kernel: 128 bytes of a lane's values
// in flight.  Eddy78, 16 complex<float> right-hand sides: 8 entries a batch
// (140 VGPRs, 3 waves per SIMD) 23.3 us per colour launch, 4 (84 VGPRs, 5
// waves) 21.0 us, 2 (58 VGPRs, 8 waves) 21.2 us with slower small colours;
// complex<double>, 2 entries, against 4 and 1: 49.2 against 50.4 and 51.1 us
// in the forward colours.
template <typename P, int W> constexpr int split_chunk()
{
    return chunk_of<P, W / split_lanes<P, W>(), 128>();
}
// One-vector SpMV and sweeps load the matrix / factor entries with the
// non-temporal hint when they stream more than this (a quarter of the 256 MB
// last-level cache): eddy78 (356 MB of A, 192 MB of factor a sweep) one-
// vector COCR 0.472 -> 0.423 s, the 64-system batch 0.507 -> 0.483 s; the
// 48^3 Laplacian (9 MB, cached from iteration to iteration) 0.0109 ->
// 0.0115 s and its SGS-MRTR 0.049 -> 0.057 s with the hint.
constexpr size_t kStreamNtBytes = size_t(64) << 20;
// right-hand sides one launch carries (larger blocks are split)
constexpr int kMaxRhs = 16;
// the block width that carries nrhs right-hand sides: 1, 4, 8 or 16
inline int rhs_width(int nrhs) { return nrhs <= 1 ? 1 : nrhs <= 4 ? 4 : nrhs <= 8 ? 8 : 16; }
static_assert(sparsesolv_amd::kSweepChunk >= 16, "layout padding shorter than a batch");
// Krylov iterations per captured graph
constexpr int kGraphIterations = 8;
// mixed precision: refinement passes per solve.  Mixed-precision COCR
// decides at the end of every captured group (kGraphIterations) whether to
// recompute rt = M^{-1} r (residual replacement): when ||r|| has fallen
// kReplaceDrop-fold since the last replacement, or kReplaceGap iterations
// have passed.  A fixed interval was measured first on the 16-RHS 78^3 eddy
// blocks: every 8 / 16 / 32 / 64 iterations the lossy system converges in
// 471-517 / 475-499 / 470-500 / 471-510 iterations (1.547 / 1.474 / 1.456 /
// 1.459 s), the configured one runs 1000 iterations in 2.749 / 2.691 /
// 2.658 / 2.640 s -- but every 32 leaves a fast-converging system (a
// complex-shifted 12^3 Laplacian) unconverged: its residual falls by
// orders of magnitude between replacements, and the recurrence's float
// error grows against the shrinking rt.
constexpr int kRefinePasses = 8;
constexpr int kReplaceGap = 32;
constexpr double kReplaceDrop = 1e-2;

inline int vec_grid(index_t n)
{
    return int(std::max<index_t>(1, std::min<index_t>((n + ssa::kBlock - 1) / ssa::kBlock, kMaxGrid)));
}


// ===========================================================================
// IC handle
// ===========================================================================

using ssa::SolverState;
using ssa::SolveParams;

template <typename T> struct HostOf { using type = T; };
template <> struct HostOf<ssa::c128> { using type = std::complex<double>; };
template <> struct HostOf<ssa::c64> { using type = std::complex<float>; };

struct DevSweep {
    DevBuf lrow, linfo, val, col;
    DevBuf slice_sys, walk_sys;   // batches: system of every slice / oversized block
    bool batched = false;
    DevBuf big_pos, big_lvl_off, big_lvl, big_rows, big_rp, big_ci, big_v;
    DevBuf w_step_off, w_row, w_ext_rp, w_ext_ci, w_ext_v, w_rec, w_sa, w_sb;
    std::vector<index_t> phase_slices, phase_big, phase_rows, phase_slot, phase_stride;
    std::vector<int64_t> phase_entries, phase_ext_rows;
    int BR = 1;
    int walk_kmax = 0;
    std::vector<index_t> h_step_off;
    size_t walk_slots = 0;   // walk slots of one right-hand side
    bool walk_checked = false, walk_fits = false;
    index_t nphases = 0, max_big_rows = 0, max_levels = 0;
    size_t stream_bytes = 0;
    // The straggler walk's place in the schedule: the one phase with
    // oversized blocks (-1: none, several, or no walk), the first later
    // phase whose slices read any of their rows (nphases: none), whether
    // those rows have entries outside their blocks, and the walk's steps.
    index_t walk_phase = -1, walk_reader = 0;
    bool walk_ext = false;
    size_t walk_steps = 0;

    template <typename F>
    void upload(const sparsesolv_amd::SweepLayout<F>& L, hipStream_t s)
    {
        lrow.upload(L.lane_row, s);
        linfo.upload(L.lane_info, s);
        val.upload(L.val, s);
        col.upload(L.col, s);
        big_pos.upload(L.big_pos, s);
        big_lvl_off.upload(L.big_lvl_off, s);
        big_lvl.upload(L.big_lvl, s);
        big_rows.upload(L.big_rows, s);
        big_rp.upload(L.big_rp, s);
        big_ci.upload(L.big_ci, s);
        big_v.upload(L.big_v, s);
        w_step_off.upload(L.walk_step_off, s);
        w_row.upload(L.walk_row, s);
        w_ext_rp.upload(L.walk_ext_rp, s);
        w_ext_ci.upload(L.walk_ext_ci, s);
        w_ext_v.upload(L.walk_ext_v, s);
        w_rec.upload(L.walk_rec, s);
        walk_kmax = L.walk_kmax;
        h_step_off = L.walk_step_off;
        walk_slots = L.walk_row.size();
        walk_scratch(1, s);
        phase_slices = L.phase_slices;
        phase_slot = L.phase_slot;
        phase_stride = L.phase_stride;
        phase_big = L.phase_big;
        phase_rows = L.phase_rows;
        phase_entries = L.phase_entries;
        phase_ext_rows = L.phase_ext_rows;
        BR = L.BR;
        nphases = L.nphases;
        max_big_rows = L.max_big_rows;
        max_levels = L.max_levels;
        stream_bytes = L.val.size() * sizeof(F) + L.col.size() * 4 + L.big_v.size() * sizeof(F) +
                       L.big_ci.size() * 4;
        place_walk(L);
    }

    template <typename F>
    void place_walk(const sparsesolv_amd::SweepLayout<F>& L)
    {
        using Lay = sparsesolv_amd::SweepLayout<F>;
        walk_phase = -1;
        walk_reader = L.nphases;
        walk_ext = false;
        walk_steps = 0;
        for (index_t q = 0; q < L.nphases; ++q) {
            if (L.phase_big[q + 1] == L.phase_big[q]) continue;
            if (walk_phase >= 0) {
                walk_phase = -1;   // several phases with oversized blocks
                return;
            }
            walk_phase = q;
        }
        if (walk_phase < 0 || L.walk_kmax == 0) {
            walk_phase = -1;
            return;
        }
        const size_t t0 = size_t(L.walk_step_off[L.phase_big[walk_phase]]) * 64;
        const size_t t1 = size_t(L.walk_step_off[L.phase_big[walk_phase + 1]]) * 64;
        walk_steps = (t1 - t0) / 64;
        walk_ext = L.walk_ext_rp[t1] != L.walk_ext_rp[t0];
        index_t maxr = -1;
        for (size_t t = t0; t < t1; ++t) maxr = std::max(maxr, L.walk_row[t]);
        std::vector<char> big(size_t(maxr + 1), 0);
        for (size_t t = t0; t < t1; ++t)
            if (L.walk_row[t] >= 0) big[size_t(L.walk_row[t])] = 1;
        for (index_t q = walk_phase + 1; q < L.nphases; ++q)
            for (index_t sl = L.phase_slices[q]; sl < L.phase_slices[q + 1]; ++sl) {
                const index_t K = Lay::info_len(L.lane_info[size_t(sl) * 64]);
                const size_t a = size_t(L.slice_slot[sl]) * 64, e = a + size_t(K) * 64;
                for (size_t k = a; k < e; ++k) {
                    const index_t c = L.col[k];
                    if (c >= 0 && c <= maxr && big[size_t(c)]) {
                        walk_reader = q;
                        return;
                    }
                }
            }
    }
    // the walk's per-slot scratch for w right-hand sides
    void walk_scratch(int w, hipStream_t s)
    {
        const size_t b = size_t(w) * walk_slots * sizeof(ssa::c128);
        // idle walk lanes read these slots and never have them written
        if (w_sa.reserve(b)) HIP_CHECK(hipMemsetAsync(w_sa.p, 0, w_sa.bytes, s));
        if (w_sb.reserve(b)) HIP_CHECK(hipMemsetAsync(w_sb.p, 0, w_sb.bytes, s));
    }
    size_t bytes() const
    {
        return lrow.bytes + linfo.bytes + val.bytes + col.bytes +
               big_pos.bytes + big_lvl_off.bytes + big_lvl.bytes + big_rows.bytes +
               big_rp.bytes + big_ci.bytes + big_v.bytes + w_step_off.bytes + w_row.bytes +
               w_ext_rp.bytes + w_ext_ci.bytes + w_ext_v.bytes + w_rec.bytes;
    }
};


template <typename T, typename F>
struct IcImpl final : IcBase {
    using HT = typename HostOf<T>::type;
    using HF = typename HostOf<F>::type;
    using P = F;  // preconditioner-internal vector type

    ssa_context* ctx;
    // count systems; n = device rows (the systems' rows, each padded to a
    // multiple of 256 when count > 1); tot = rows of the caller's vectors
    int count = 1;
    index_t n = 0, tot = 0, nnz = 0, nnz_l = 0, nc = 0, nb = 0;
    double actual_shift = 0.0, setup_seconds = 0.0;
    int shift_trials = 0;
    // host copies kept for queries (one system)
    std::vector<index_t> h_ordering, h_color_offsets, h_block_offsets, h_Lrp, h_Lci;
    std::vector<HT> h_Lv, h_invd;
    // batches: per system row ranges and grid ranges
    std::vector<int> h_row0, h_nrows, h_vec_g0, h_spmv_g0;
    DevBuf d_vec_sys, d_vec_g0, d_spmv_sys, d_spmv_g0, d_row0, d_nrows;
    int vec_total = 0;
    int hist_stride = 0;

    // device: operator, permutations, factor
    DevBuf A_off, A_val, A_col, A_row, perm, rev, invd;
    index_t A_nslices = 0;
    size_t spmv_bytes = 0;
    DevSweep fwd, bwd;
    // device: work space
    DevBuf r, z, p, Ap, t, qt, x, best, bvec, tmp, tmp2;
    DevBuf part_a, part_b, part_c, part_d, state, hist;
    DevBuf rstate;   // mixed COCR: which states recompute rt (status 0)
    DevBuf idle;   // states that never stop: sweeps outside a solve
    // right-hand sides of the current operation (every vector holds W blocks
    // of n rows; W > 1 for one system only) and the width the work space
    // was sized for
    int W = 1, ws_width = 0;
    DevBuf pass;                          // PassParams per state
    ssa::PassParams* h_pass = nullptr;    // pinned staging for `pass`
    int h_cap = 0;                        // states h_state / h_pass hold
    // Mixed precision (P narrower than T): the preconditioner's input, output
    // and intermediate vectors in P.  In FP64 they alias the Krylov vectors.
    static constexpr bool kMixed = !std::is_same<T, P>::value;
    // steps the straggler walk keeps in flight: the compiler waits for all
    // of them at the start of each group, so deeper groups pay that round
    // trip less often; complex<double> operands, and rows of 8 internal
    // entries, take too many registers for 16 (those walks prefetch 8)
    static constexpr int kWalkPrefetch = sizeof(P) > 8 ? 8 : 16;
    // The straggler walks beside the colours that do not read them: a
    // second stream (default priority) forks after the forward colour kernel
    // of the walk's phase and joins before the first backward phase that
    // reads the walk's rows.  A fork and join cost ~14 us on the main
    // stream, so by default (walk_overlap = -1) it is used when the walks
    // take at least kOverlapSteps steps (~0.13 us each).
    static constexpr size_t kOverlapSteps = 160;
    int overlap_setting = -1;
    int overlap = -1;   // -1 undecided, 0 serial, 1 overlapped
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipStream_t sweep_stream = nullptr;   // where sweep kernels go (null: the context's)
    bool serial_sweeps = false;           // no walk overlap (no second stream is made)
    static_assert(sparsesolv_amd::kWalkDepth % kWalkPrefetch == 0, "walk padding");
    // mixed: q (lo_a, the sweeps' input), qt (lo_b) and the sweeps'
    // intermediate (lo_t) in P, and COCR's rt = M^-1 r (lo_r): rt comes from
    // the P preconditioner, is updated by rt -= alpha qt with qt in P and is
    // the SpMV's input t = A rt, so it is kept in P (computed in T, stored
    // rounded): the SpMV gathers and the update move half the bytes for it
    DevBuf lo_a, lo_b, lo_t, lo_r;
    // diagonal_scaling: S = diag(1/sqrt|a_ii|) per device row and the
    // preconditioner's scaled input.  The reference applies S M^-1 S on the
    // level path and in ICCG's reordered-space apply (Path 1,
    // abmc_reorder_spmv), and M^-1 of the scaled factor unscaled in the
    // ABMC apply (the composite scaling vector it uses is built before the
    // factors, ic_preconditioner.hpp:125 vs :162); scale_apply holds the
    // current operation's choice.
    DevBuf scale_d, sbuf;
    bool has_scale = false, abmc_order = true, abmc_rcm = false, scale_apply = false;
    // what the factor is (sparsesolv_amd::kFactorIc / kFactorSgs /
    // kFactorSgsMrtr); SGS-MRTR also keeps L = lower(D A D) as CSR (for its
    // normaliser |L D b|) and D, D^-1, all in device order
    int factor_kind = sparsesolv_amd::kFactorIc;
    int rhs_block = 0;   // right-hand sides per launch (0: kMaxRhs)
    DevBuf mr_rp, mr_ci, mr_v, mr_d, mr_id;

    SolverState* h_state = nullptr;  // pinned, one per state

    // Kernel timing inside solves (ssa_ic_trace): while a kernel class is
    // selected, solves run without graph capture and each launch of that
    // class in the iteration loop is bracketed by two events on the stream
    // it runs on; the pairs are read after every status poll (the stream
    // has drained by then) and summed with the launch's algorithmic bytes.
    int trace_cls = -1;
    bool trace_armed = false;   // inside a solve's iteration loop
    struct TracePair {
        hipEvent_t a = nullptr, b = nullptr;
        double bytes = 0.0;
    };
    std::vector<TracePair> trace_pool;
    size_t trace_used = 0;
    long long trace_launches = 0;
    double trace_us = 0.0, trace_bytes = 0.0;

    template <typename Fn>
    void traced(int cls, hipStream_t s, double bytes, Fn&& launch)
    {
        if (!trace_armed || cls != trace_cls) {
            launch();
            return;
        }
        if (trace_used == trace_pool.size()) {
            // timing-only events: no system-scope fence (cache writeback and
            // invalidation) at record, which would lengthen the bracketed
            // launch; they are read after the stream has been synchronised
            TracePair tp;
            HIP_CHECK(hipEventCreateWithFlags(&tp.a, hipEventDisableSystemFence));
            if (hipEventCreateWithFlags(&tp.b, hipEventDisableSystemFence) != hipSuccess) {
                (void)hipEventDestroy(tp.a);
                throw SsaError(SSA_ERR_DEVICE, "ssa_ic_trace: hipEventCreate failed");
            }
            trace_pool.push_back(tp);
        }
        TracePair& tp = trace_pool[trace_used++];
        tp.bytes = bytes;
        HIP_CHECK(hipEventRecord(tp.a, s));
        launch();
        HIP_CHECK(hipEventRecord(tp.b, s));
    }

    // the recorded launches into the totals (the streams have drained)
    void trace_harvest()
    {
        for (size_t i = 0; i < trace_used; ++i) {
            float ms = 0.f;
            HIP_CHECK(hipEventElapsedTime(&ms, trace_pool[i].a, trace_pool[i].b));
            trace_us += 1000.0 * double(ms);
            trace_bytes += trace_pool[i].bytes;
            ++trace_launches;
        }
        trace_used = 0;
    }

    void trace(int what) override
    {
        need_device();
        if (what < -1 || what > SSA_TRACE_DOT) throw SsaError(SSA_ERR_ARGUMENT, "ssa_ic_trace: unknown kernel class");
        trace_cls = what;
        trace_used = 0;
        if (what >= 0) {   // a new selection starts from zero; off keeps the totals
            trace_launches = 0;
            trace_us = trace_bytes = 0.0;
        }
    }

    void trace_read(long long& launches, double& us, double& bytes) const override
    {
        launches = trace_launches;
        us = trace_us;
        bytes = trace_bytes;
    }

    // algorithmic bytes of one launch of a vector kernel class: every
    // operand once, per right-hand side (the best-x copy is not counted)
    double vector_bytes(int cls, bool cocr) const
    {
        const double e = double(n) * double(W), sv = double(sizeof(T)), sp = double(sizeof(P));
        const double low = kMixed ? sp : 0.0;   // the narrow copy a mixed kernel also writes
        const double sq = kMixed ? sp : sv;    // qt, z, COCR's rt: the preconditioner's precision
        switch (cls) {
        case SSA_TRACE_UPDATE:
            // COCR: x, p, r, q, rt, qt read; x, r, rt (blocks: and p) written
            if (cocr) return e * ((W > 1 ? 7.0 : 6.0) * sv + 3.0 * sq);
            return e * (6.0 * sv + low);       // CG: x, p, r, Ap read; x, r written
        case SSA_TRACE_PQ:
            // COCR: t, q read, q written (one vector: also rt, p read, p written)
            if (cocr) return e * (W > 1 ? 3.0 * sv + low : 5.0 * sv + sq + low);
            return e * (2.0 * sv + sq);        // CG: z, p read; p written
        default:                 // dot: two operands
            return e * (sv + (kMixed ? sp : sv));
        }
    }

    IcImpl(ssa_context* c, int count_, const int* ns, const int32_t* const* rp,
           const int32_t* const* ci, const HT* const* v, const sparsesolv_amd::IcSettings& st)
        : ctx(c), count(count_), abmc_order(st.use_abmc), abmc_rcm(st.use_abmc && st.abmc_use_rcm)
    {
        overlap_setting = st.walk_overlap;
        factor_kind = st.factor_kind;
        rhs_block = st.rhs_block;
        ssa_host::BatchPlan<HT, HF> plan;
        ssa_host::build_batch<HT, HF>(plan, count, ns, rp, ci, v, st);
        n = plan.row0.back();
        tot = plan.in_off.back();
        setup_seconds = plan.wall_seconds;
        for (int s = 0; s < count; ++s) {
            nnz += plan.nnz[s];
            nnz_l += plan.nnz_l[s];
            nc = std::max<index_t>(nc, plan.num_colors[s]);
            nb += plan.num_blocks[s];
            actual_shift = std::max(actual_shift, plan.actual_shift[s]);
            shift_trials = std::max(shift_trials, plan.shift_trials[s]);
        }
        h_row0 = plan.row0;
        h_nrows = plan.nrows;
        if (count == 1) {
            // the plan is local: its schedule and factor move to the handle
            auto& f = plan.single.fac;
            h_ordering = std::move(f.sched.ordering);
            h_color_offsets = std::move(f.sched.color_offsets);
            h_block_offsets = std::move(f.sched.block_offsets);
            h_Lrp = std::move(f.L.rp);
            h_Lci = std::move(f.L.ci);
            h_Lv = std::move(f.L.v);
            h_invd = std::move(f.inv_diag);
        }

        if (!ctx) return;   // host-only handle: schedule and factor queries only
        hipStream_t s = ctx->stream;
        A_off.upload(plan.spmv.slice_off, s);
        A_val.upload(plan.spmv.val, s);
        A_col.upload(plan.spmv.col, s);
        A_row.upload(plan.spmv.row, s);
        A_nslices = plan.spmv.nslices;
        spmv_bytes = plan.spmv.val.size() * sizeof(HT) + plan.spmv.col.size() * 4;
        perm.upload(plan.gather_out, s);
        rev.upload(plan.gather_in, s);
        invd.upload(plan.invd, s);
        if (factor_kind == sparsesolv_amd::kFactorSgsMrtr) {
            mr_rp.upload(plan.lcsr.rp, s);
            mr_ci.upload(plan.lcsr.ci, s);
            mr_v.upload(plan.lcsr.v, s);
            mr_d.upload(plan.split_d, s);
            mr_id.upload(plan.split_id, s);
        }
        if (!plan.scale.empty()) {
            scale_d.upload(plan.scale, s);
            has_scale = true;
        }
        fwd.upload(plan.fwd, s);
        bwd.upload(plan.bwd, s);
        // the host packs walk records for values of sizeof(HF) bytes, the
        // kernel reads them as F: same size, same layout
        static_assert(sizeof(HF) == sizeof(F), "walk record value size");
        for (const auto* L : {&plan.fwd, &plan.bwd}) {
            const int K = L->walk_kmax;
            const int expect = K == 2 ? ssa::walk_words<F, 2>() : K == 4 ? ssa::walk_words<F, 4>()
                             : K == 8 ? ssa::walk_words<F, 8>() : 0;
            if (L->walk_rec_words != expect)
                throw SsaError(SSA_ERR_RUNTIME, "walk record layout mismatch");
        }
        if (count > 1) {
            // each system keeps the grid a one-system solve would use, so its
            // reductions give the same bits
            std::vector<int> vsys, ssys;
            h_vec_g0.assign(1, 0);
            h_spmv_g0.assign(1, 0);
            for (int q = 0; q < count; ++q) {
                const int g = vec_grid(plan.nrows[q]);
                for (int k = 0; k < g; ++k) vsys.push_back(q);
                h_vec_g0.push_back(h_vec_g0.back() + g);
                const int gs = (plan.row0[q + 1] - plan.row0[q]) / (64 * ssa::kWavesPerBlock);
                for (int k = 0; k < gs; ++k) ssys.push_back(q);
                h_spmv_g0.push_back(h_spmv_g0.back() + gs);
            }
            vec_total = h_vec_g0.back();
            d_vec_sys.upload(vsys, s);
            d_vec_g0.upload(h_vec_g0, s);
            d_spmv_sys.upload(ssys, s);
            d_spmv_g0.upload(h_spmv_g0, s);
            d_row0.upload(plan.row0, s);
            d_nrows.upload(plan.nrows, s);
            fwd.slice_sys.upload(plan.slice_sys_f, s);
            bwd.slice_sys.upload(plan.slice_sys_b, s);
            fwd.walk_sys.upload(plan.walk_sys_f, s);
            bwd.walk_sys.upload(plan.walk_sys_b, s);
            fwd.batched = bwd.batched = true;
            // oversized blocks must take the walk: the level kernel's arrays
            // are not combined
            for (DevSweep* D : {&fwd, &bwd})
                if (D->phase_big.back() > 0 && !walk_lds_ok(*D))
                    throw SsaError(SSA_ERR_UNSUPPORTED,
                                   "batch: an oversized block does not fit the walk's LDS");
        }
        HIP_CHECK(hipStreamSynchronize(s));   // host plan is freed on return
    }

    // the systems descriptor kernels take (null tables for one system)
    ssa::Systems systems() const
    {
        if (count == 1)
            return ssa::Systems{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, hist_stride};
        return ssa::Systems{d_vec_sys.as<int>(), d_vec_g0.as<int>(), d_spmv_sys.as<int>(),
                            d_spmv_g0.as<int>(), d_row0.as<int>(), d_nrows.as<int>(), hist_stride};
    }
    // vector-grid workgroups per right-hand side, and the launch grid
    int vgrid() const { return count == 1 ? vec_grid(n) : vec_total; }
    // solver states: one per system of a batch, one per right-hand side
    int nstates() const { return count > 1 ? count : W; }
    int system_count() const override { return count; }

    void need_device() const
    {
        if (!ctx)
            throw SsaError(SSA_ERR_ARGUMENT, "operation needs a handle created with a device context");
    }

    ~IcImpl() override
    {
        if (h_state) (void)hipHostFree(h_state);
        if (h_pass) (void)hipHostFree(h_pass);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (side) (void)hipStreamDestroy(side);
        for (auto& tp : trace_pool) {
            (void)hipEventDestroy(tp.a);
            (void)hipEventDestroy(tp.b);
        }
    }

    // ---- building blocks ---------------------------------------------------
    // work space for `width` right-hand sides; sets W
    void ensure_workspace(int width = 1)
    {
        need_device();
        if (width > 1 && count > 1)
            throw SsaError(SSA_ERR_UNSUPPORTED, "several right-hand sides need a one-system handle");
        W = std::max(width, 1);
        if (W == ws_width) return;
        hipStream_t s = ctx->stream;
        const int nst = nstates();
        if (nst > h_cap) {
            if (h_state) (void)hipHostFree(h_state);
            if (h_pass) (void)hipHostFree(h_pass);
            h_state = nullptr;
            h_pass = nullptr;
            h_cap = 0;
            HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&h_state), sizeof(SolverState) * nst));
            HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&h_pass), sizeof(ssa::PassParams) * nst));
            h_cap = nst;
        }
        const size_t vb = size_t(n) * size_t(W) * sizeof(T);
        // zeroed once: the padding rows between the systems of a batch are
        // never written and must stay zero (the SpMV's fused dot multiplies
        // them by zero entries)
        for (DevBuf* b : {&r, &z, &p, &Ap, &t, &qt, &x, &best, &bvec, &tmp, &tmp2})
            if (b->reserve(vb)) HIP_CHECK(hipMemsetAsync(b->p, 0, vb, s));
        const size_t np = size_t(std::max<index_t>(std::max(kMaxGrid, vec_total), (A_nslices + 3) / 4)) *
                          size_t(W) * sizeof(ssa::c128);
        for (DevBuf* b : {&part_a, &part_b, &part_c, &part_d}) b->reserve(np);
        state.reserve(sizeof(SolverState) * nst);
        rstate.reserve(sizeof(SolverState) * nst);
        pass.reserve(sizeof(ssa::PassParams) * nst);
        if (kMixed)
            for (DevBuf* b : {&lo_a, &lo_b, &lo_t, &lo_r})
                if (b->reserve(size_t(n) * size_t(W) * sizeof(P)))
                    HIP_CHECK(hipMemsetAsync(b->p, 0, b->bytes, s));
        if (has_scale) sbuf.reserve(size_t(n) * size_t(W) * sizeof(P));
        if (idle.reserve(sizeof(SolverState) * nst)) HIP_CHECK(hipMemsetAsync(idle.p, 0, idle.bytes, s));
        fwd.walk_scratch(W, s);
        bwd.walk_scratch(W, s);
        ws_width = W;
    }

    int spmv_grid() const { return int((A_nslices + ssa::kWavesPerBlock - 1) / ssa::kWavesPerBlock); }

    // y = A x (ABMC order) for the W right-hand sides, optional fused dot
    // into part
    void launch_spmv(const T* xin, T* y, int dot, T* part, const SolverState* st,
                     const ssa::Ctrl& ctrl = ssa::Ctrl{})
    {
        switch (W) {
        case 1: spmv_w<1>(xin, y, dot, part, st, ctrl); break;
        case 4: spmv_w<4>(xin, y, dot, part, st, ctrl); break;
        case 8: spmv_w<8>(xin, y, dot, part, st, ctrl); break;
        case 16: spmv_w<16>(xin, y, dot, part, st, ctrl); break;
        default: throw SsaError(SSA_ERR_ARGUMENT, "unsupported block width");
        }
        HIP_CHECK(hipGetLastError());
    }

    // y = A x with x in the preconditioner's precision (widened exactly)
    void launch_spmv_low(const P* xin, T* y, int dot, T* part, const SolverState* st,
                         const ssa::Ctrl& ctrl = ssa::Ctrl{})
    {
        switch (W) {
        case 1: spmv_w<1, P>(xin, y, dot, part, st, ctrl); break;
        case 4: spmv_w<4, P>(xin, y, dot, part, st, ctrl); break;
        case 8: spmv_w<8, P>(xin, y, dot, part, st, ctrl); break;
        case 16: spmv_w<16, P>(xin, y, dot, part, st, ctrl); break;
        default: throw SsaError(SSA_ERR_ARGUMENT, "unsupported block width");
        }
        HIP_CHECK(hipGetLastError());
    }

    template <int WV, typename TX = T>
    void spmv_w(const TX* xin, T* y, int dot, T* part, const SolverState* st, const ssa::Ctrl& ctrl)
    {
        const int g = std::max(1, spmv_grid());
        hipStream_t s = ctx->stream;
        const ssa::Systems sy = systems();
        const index_t* off = A_off.as<index_t>();
        const index_t* col = A_col.as<index_t>();
        const index_t* rows = A_row.as<index_t>();
        const T* val = A_val.as<T>();
        const int nt = WV == 1 && A_val.bytes + A_col.bytes > kStreamNtBytes ? 1 : 0;
        // read x once, write y, the matrix entries and row pointers once
        const double bytes = double(nnz) * (sizeof(T) + 4.0) + 4.0 * (double(n) + 1.0) +
                             double(n) * WV * (sizeof(TX) + sizeof(T));
        traced(SSA_TRACE_SPMV, s, bytes, [&] {
            if (dot == 0)
                ssa::k_spmv<T, 0, WV, TX><<<g, ssa::kBlock, 0, s>>>(off, val, col, rows, A_nslices, xin, y, part, st, ctrl, sy, nt);
            else if (dot == 1)
                ssa::k_spmv<T, 1, WV, TX><<<g, ssa::kBlock, 0, s>>>(off, val, col, rows, A_nslices, xin, y, part, st, ctrl, sy, nt);
            else
                ssa::k_spmv<T, 2, WV, TX><<<g, ssa::kBlock, 0, s>>>(off, val, col, rows, A_nslices, xin, y, part, st, ctrl, sy, nt);
        });
    }

    ssa::SweepArrays sweep_arrays(DevSweep& D)
    {
        return ssa::SweepArrays{D.lrow.as<index_t>(), D.linfo.as<index_t>(),
                                D.val.p, D.col.as<index_t>(), invd.p,
                                D.batched ? D.slice_sys.as<int>() : nullptr,
                                W == 1 && D.val.bytes + D.col.bytes > kStreamNtBytes ? 1 : 0};
    }

    // a colour small enough to share a single-workgroup kernel with its
    // neighbours: no oversized block, at most two passes of the workgroup
    int tail_threads(int BR) const
    {
        if (W > 1) return BR >= 16 ? ssa::tail_threads<P, 16, 4>() : ssa::tail_threads<P, 1, 4>();
        return BR >= 16 ? ssa::tail_threads<P, 16>() : ssa::tail_threads<P, 1>();
    }

    // A run of small colours shares one workgroup, i.e. one CU's address
    // rate: with several right-hand sides (a cache line per lane and entry)
    // a colour joins only while its gathers are few -- the eddy78 16-vector
    // forward tail, 0.8 MB of gathers, took 39 us in one workgroup and 21
    // us as two launches of a wavefront per workgroup; the backward one, 0.1
    // MB, 15 us against 19.
    static constexpr double kTailGatherBytes = 128.0 * 1024.0;
    bool small_phase(const DevSweep& D, index_t ph) const
    {
        return (W == 1 || double(D.phase_entries[ph]) * W * sizeof(P) <= kTailGatherBytes) &&
               D.phase_big[ph + 1] == D.phase_big[ph] &&
               D.phase_slices[ph + 1] - D.phase_slices[ph] <= 2 * (tail_threads(D.BR) / 64);
    }

    // end of the run of small colours that starts at ph (ph itself when
    // fewer than two colours would share the tail kernel)
    index_t tail_end(const DevSweep& D, index_t ph, index_t limit = -1) const
    {
        if (limit < 0) limit = D.nphases;
        index_t e = ph;
        while (e < limit && e - ph < ssa::kTailMaxPhases && small_phase(D, e)) ++e;
        return e - ph >= 2 ? e : ph;
    }

    bool in_tail_group(const DevSweep& D, index_t ph) const
    {
        for (index_t q = 0; q < D.nphases;) {
            const index_t e = tail_end(D, q);
            if (e > q) {
                if (ph >= q && ph < e) return true;
                q = e;
            } else {
                ++q;
            }
        }
        return false;
    }

    // what launch_phase enqueues of a phase: everything; the colour kernel
    // with the external part of the oversized blocks but not their walk;
    // the regular slices only; the external part only; the walk only
    enum PhasePart { kWhole, kNoWalk, kRegularOnly, kExtOnly, kWalkOnly };

    hipStream_t lstream() const { return sweep_stream ? sweep_stream : ctx->stream; }

    // all colours of one sweep; runs of small colours go to one workgroup
    template <bool FWD>
    void launch_sweep(DevSweep& D, const P* in, P* out, const SolverState* st)
    {
        launch_phases<FWD>(D, 0, D.nphases, in, out, st);
    }

    // phases [a, b); phase `special` (never in a run of small colours: it
    // has oversized blocks) enqueues only `part`
    template <bool FWD>
    void launch_phases(DevSweep& D, index_t a, index_t b, const P* in, P* out,
                       const SolverState* st, index_t special = -1, PhasePart part = kWhole)
    {
        index_t ph = a;
        while (ph < b) {
            const index_t e = tail_end(D, ph, b);
            if (e > ph) {
                launch_tail<FWD>(D, ph, e, in, out, st);
                ph = e;
            } else {
                launch_phase<FWD>(D, ph, in, out, st, ph == special ? part : kWhole);
                ++ph;
            }
        }
    }

    // launches of phases [a, b) as launch_phases groups them
    index_t count_launches(const DevSweep& D, index_t a, index_t b) const
    {
        index_t k = 0;
        for (index_t ph = a; ph < b; ++k) {
            const index_t e = tail_end(D, ph, b);
            ph = e > ph ? e : ph + 1;
        }
        return k;
    }

    // whether the walks run on the second stream; decided (and the stream
    // made) at the first apply
    bool use_overlap()
    {
        if (overlap >= 0) return overlap == 1;
        overlap = 0;
        const bool shape = overlap_setting != 0 && fwd.walk_phase >= 0 && bwd.walk_phase >= 0 &&
                           fwd.walk_reader == fwd.nphases && !bwd.walk_ext &&
                           walk_lds_ok(fwd) && walk_lds_ok(bwd);
        if (!shape) return false;
        const index_t window = count_launches(fwd, fwd.walk_phase + 1, fwd.nphases) +
                               count_launches(bwd, 0, bwd.walk_reader);
        if (overlap_setting < 0 && (fwd.walk_steps + bwd.walk_steps < kOverlapSteps || window < 3))
            return false;
        HIP_CHECK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
        overlap = 1;
        return true;
    }

    // Both sweeps, out = M^-1 in through tbuf.  Overlapped: the forward
    // walk reads the external sums its colour kernel wrote and writes rows
    // no later forward colour reads; the backward walk's rows read no row
    // outside their blocks (no external part), so its input -- the forward
    // walk's rows of tbuf -- is all it waits for; the main stream joins
    // before the first backward phase that reads the walk's rows.
    void launch_sweeps(const P* in, P* out, P* tbuf, const SolverState* st)
    {
        if (serial_sweeps || !use_overlap()) {
            launch_sweep<true>(fwd, in, tbuf, st);
            launch_sweep<false>(bwd, tbuf, out, st);
            return;
        }
        const index_t pf = fwd.walk_phase, pb = bwd.walk_phase, qb = bwd.walk_reader;
        hipStream_t s = ctx->stream;
        launch_phases<true>(fwd, 0, pf + 1, in, tbuf, st, pf, kNoWalk);
        HIP_CHECK(hipEventRecord(ev_fork, s));
        HIP_CHECK(hipStreamWaitEvent(side, ev_fork, 0));
        struct Restore {
            hipStream_t& v;
            ~Restore() { v = nullptr; }
        } restore{sweep_stream};
        sweep_stream = side;
        launch_phase<true>(fwd, pf, in, tbuf, st, kWalkOnly);
        launch_phase<false>(bwd, pb, tbuf, out, st, kExtOnly);
        launch_phase<false>(bwd, pb, tbuf, out, st, kWalkOnly);
        sweep_stream = nullptr;
        HIP_CHECK(hipEventRecord(ev_join, side));
        launch_phases<true>(fwd, pf + 1, fwd.nphases, in, tbuf, st);
        launch_phases<false>(bwd, 0, qb, tbuf, out, st, pb, kRegularOnly);
        HIP_CHECK(hipStreamWaitEvent(s, ev_join, 0));
        launch_phases<false>(bwd, qb, bwd.nphases, tbuf, out, st);
    }

    template <bool FWD>
    void launch_tail(DevSweep& D, index_t p0, index_t p1, const P* in, P* out, const SolverState* st)
    {
        if (!st) st = idle.as<SolverState>();
        ssa::TailPhases tp{};
        tp.count = int(p1 - p0);
        for (index_t q = p0; q <= p1; ++q) tp.slices[q - p0] = D.phase_slices[q];
        for (index_t q = p0; q < p1; ++q) {
            tp.slot[q - p0] = D.phase_slot[q];
            tp.stride[q - p0] = D.phase_stride[q];
        }
        const ssa::SweepArrays A = sweep_arrays(D);
#define SSA_TAIL(BRV) tail_br<FWD, BRV>(tp, A, in, out, st)
        switch (D.BR == 0 ? 1 : D.BR) {
        case 1: SSA_TAIL(1); break;
        case 2: SSA_TAIL(2); break;
        case 4: SSA_TAIL(4); break;
        case 8: SSA_TAIL(8); break;
        case 16: SSA_TAIL(16); break;
        default: throw SsaError(SSA_ERR_UNSUPPORTED, "sweep: unsupported rows per lane");
        }
#undef SSA_TAIL
        HIP_CHECK(hipGetLastError());
    }

    // a run of small colours in one workgroup, WV right-hand sides per lane
    template <bool FWD, int BRV, int WV>
    void tail_w(const ssa::TailPhases& tp, const ssa::SweepArrays& A, const P* in, P* out,
                const SolverState* st)
    {
        // one workgroup: batches as wide as the register budget of its
        // waves allows (256 threads: a wave a SIMD, as in small colours;
        // 512: two, 512 bytes a lane)
        constexpr int TT = ssa::tail_threads<P, BRV, WV>();
        constexpr int CH = WV == 1 ? 8 : TT > 256 ? chunk_of<P, WV, 512>() : sweep_chunk_small<P, WV>();
        ssa::k_sweep_tail<P, F, BRV, FWD, CH, WV><<<1, ssa::tail_threads<P, BRV, WV>(), 0, lstream()>>>(
            tp, A, in, out, st);
    }

    template <bool FWD, int BRV>
    void tail_br(const ssa::TailPhases& tp, const ssa::SweepArrays& A, const P* in, P* out,
                 const SolverState* st)
    {
        switch (W) {
        case 1: tail_w<FWD, BRV, 1>(tp, A, in, out, st); break;
        case 4: tail_w<FWD, BRV, 4>(tp, A, in, out, st); break;
        case 8: tail_w<FWD, BRV, 8>(tp, A, in, out, st); break;
        case 16: tail_w<FWD, BRV, 16>(tp, A, in, out, st); break;
        default: throw SsaError(SSA_ERR_ARGUMENT, "unsupported block width");
        }
    }

    // one colour's regular slices (and the external part of its oversized
    // blocks, gw workgroups) with WV right-hand sides per lane.  Several
    // right-hand sides take split_lanes lanes a row (k_sweep_split, the
    // external part in workgroups after the slices': the eddy78 forward
    // colour with the oversized block 30.6 us with a lane a row, 24.2 us
    // split, 19.7 + 14.2 us as two launches); a small colour of
    // them without an external part runs a wavefront per workgroup, since
    // its gathers are bound by the address rate of the CUs they run on.
    template <bool FWD, int BRV, int WV>
    void sweep_w(int gw, index_t s0, index_t s1, index_t slot0, index_t stride,
                 const ssa::SweepArrays& A, const P* in, P* out, const ssa::WalkExt& we,
                 const SolverState* st)
    {
        const index_t ns = s1 - s0;
        if constexpr (WV > 1) {
            if (ns > 0) {
                constexpr int L = split_lanes<P, WV>();
                constexpr int CH = split_chunk<P, WV>();
                static_assert(64 / L % sparsesolv_amd::kBlockAlign == 0, "a wavefront's rows hold whole blocks");
                const size_t waves = size_t(ns) * L;
                if (gw == 0 && ns <= kSmallColourSlices)
                    ssa::k_sweep_split<P, F, BRV, FWD, CH, WV, L, 1><<<unsigned(waves), 64, 0, lstream()>>>(
                        s0, s1, slot0, stride, A, in, out, we, st);
                else
                    ssa::k_sweep_split<P, F, BRV, FWD, CH, WV, L, ssa::kWavesPerBlock>
                        <<<unsigned((waves + ssa::kWavesPerBlock - 1) / ssa::kWavesPerBlock) + unsigned(gw),
                           ssa::kBlock, 0, lstream()>>>(s0, s1, slot0, stride, A, in, out, we, st);
                return;
            }
        }
        const int g = int((ns + ssa::kWavesPerBlock - 1) / ssa::kWavesPerBlock) + gw;
        ssa::k_sweep<P, F, BRV, FWD, sweep_chunk<P, WV>(), WV><<<g, ssa::kBlock, 0, lstream()>>>(
            s0, s1, slot0, stride, A, in, out, we, st);
    }

    template <bool FWD, int BRV>
    void sweep_br(int gw, index_t s0, index_t s1, index_t slot0, index_t stride,
                  const ssa::SweepArrays& A, const P* in, P* out, const ssa::WalkExt& we,
                  const SolverState* st)
    {
        switch (W) {
        case 1: sweep_w<FWD, BRV, 1>(gw, s0, s1, slot0, stride, A, in, out, we, st); break;
        case 4: sweep_w<FWD, BRV, 4>(gw, s0, s1, slot0, stride, A, in, out, we, st); break;
        case 8: sweep_w<FWD, BRV, 8>(gw, s0, s1, slot0, stride, A, in, out, we, st); break;
        case 16: sweep_w<FWD, BRV, 16>(gw, s0, s1, slot0, stride, A, in, out, we, st); break;
        default: throw SsaError(SSA_ERR_ARGUMENT, "unsupported block width");
        }
    }

    // algorithmic bytes of colour ph's kernel: its factor entries (value and
    // index, the diagonal counted), a row pointer, the input and output
    // entry of its rows, and each distinct row of earlier colours it reads,
    // once (W values per row)
    double colour_bytes(const DevSweep& D, index_t ph) const
    {
        return double(D.phase_entries[ph]) * (sizeof(F) + 4.0) +
               double(D.phase_rows[ph]) * (4.0 + 2.0 * W * sizeof(P)) +
               double(D.phase_ext_rows[ph]) * W * sizeof(P);
    }

    template <bool FWD>
    void launch_phase(DevSweep& D, index_t ph, const P* in, P* out, const SolverState* st,
                      PhasePart part = kWhole)
    {
        if (!st) st = idle.as<SolverState>();
        hipStream_t s = lstream();
        index_t s0 = D.phase_slices[ph], s1 = D.phase_slices[ph + 1];
        const index_t b0 = D.phase_big[ph], b1 = D.phase_big[ph + 1];
        const bool walk = b1 > b0 && D.walk_kmax > 0 && walk_lds_ok(D);
        ssa::WalkExt we{};
        int gw = 0;
        if (walk) {
            we.t0 = size_t(D.h_step_off[b0]) * 64;
            we.t1 = size_t(D.h_step_off[b1]) * 64;
            we.row = D.w_row.as<index_t>();
            we.ext_rp = D.w_ext_rp.as<index_t>();
            we.ext_ci = D.w_ext_ci.as<index_t>();
            we.ext_v = D.w_ext_v.p;
            we.sa = D.w_sa.p;
            we.sb = D.w_sb.p;
            we.ws = D.walk_slots;
            gw = int((we.t1 - we.t0 + ssa::kBlock - 1) / ssa::kBlock);
        }
        if (part == kWalkOnly || part == kExtOnly) s1 = s0;   // no regular slices
        if (part == kWalkOnly || part == kRegularOnly) gw = 0;
        if (s1 > s0 || gw > 0) {
            const ssa::SweepArrays A = sweep_arrays(D);
            // timed as the colour kernel: colours without oversized blocks
            const int cls = b1 == b0 ? SSA_TRACE_COLOUR : -2;
            const double bytes = colour_bytes(D, ph);
#define SSA_SWEEP(BRV) \
    sweep_br<FWD, BRV>(gw, s0, s1, D.phase_slot[ph], D.phase_stride[ph], A, in, out, we, st)
            traced(cls, s, bytes, [&] {
                switch (D.BR == 0 ? 1 : D.BR) {
                case 1: SSA_SWEEP(1); break;
                case 2: SSA_SWEEP(2); break;
                case 4: SSA_SWEEP(4); break;
                case 8: SSA_SWEEP(8); break;
                case 16: SSA_SWEEP(16); break;
                default: throw SsaError(SSA_ERR_UNSUPPORTED, "sweep: unsupported rows per lane");
                }
            });
#undef SSA_SWEEP
            HIP_CHECK(hipGetLastError());
        }
        if (walk && (part == kWhole || part == kWalkOnly)) {
            const size_t lds = (size_t(D.max_big_rows) + 2) * sizeof(P);
#define SSA_WALK(KM, DEPTH)                                                                    \
    ssa::k_sweep_walk<P, F, FWD, KM, DEPTH><<<dim3(unsigned(b1 - b0), unsigned(W)), 64, lds, s>>>( \
        b0, D.w_step_off.as<index_t>(), D.w_rec.as<uint4>(), D.max_big_rows, out,              \
        D.w_sa.as<P>(), D.w_sb.as<P>(), W, D.walk_slots, st,                                  \
        D.batched ? D.walk_sys.as<int>() : nullptr)
            switch (D.walk_kmax) {
            case 2: SSA_WALK(2, kWalkPrefetch); break;
            case 4: SSA_WALK(4, kWalkPrefetch); break;
            default: SSA_WALK(8, 8); break;
            }
#undef SSA_WALK
            HIP_CHECK(hipGetLastError());
        } else if (!walk && b1 > b0) {
            // level fallback (never split: the overlap needs the walk)
            const size_t lds = size_t(D.max_big_rows) * sizeof(P);
            const int use_lds = lds <= 64 * 1024 ? 1 : 0;
            ssa::k_sweep_big<P, F, FWD><<<dim3(unsigned(b1 - b0), unsigned(W)), 1024, use_lds ? lds : 0, s>>>(
                b0, D.big_pos.as<index_t>(), D.big_lvl_off.as<index_t>(), D.big_lvl.as<index_t>(),
                D.big_rows.as<index_t>(), D.big_rp.as<index_t>(), D.big_ci.as<index_t>(),
                D.big_v.as<F>(), invd.as<F>(), in, out, use_lds, W, st);
            HIP_CHECK(hipGetLastError());
        }
    }

    // whether the walk's LDS (one value per big-block row + a zero slot)
    // fits the workgroup limit; raises the kernels' dynamic LDS cap once
    bool walk_lds_ok(DevSweep& D)
    {
        const size_t need = (size_t(D.max_big_rows) + 2) * sizeof(P);
        if (D.walk_checked) return D.walk_fits;
        D.walk_checked = true;
        int optin = 0;
        if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, ctx->device) != hipSuccess ||
            optin <= 0)
            HIP_CHECK(hipDeviceGetAttribute(&optin, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
        D.walk_fits = need <= size_t(optin);
        if (D.walk_fits && need > 64 * 1024) {
            const int bytes = int(need);
            HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&ssa::k_sweep_walk<P, F, true, 2, kWalkPrefetch>), hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
            HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&ssa::k_sweep_walk<P, F, true, 4, kWalkPrefetch>), hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
            HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&ssa::k_sweep_walk<P, F, true, 8, 8>), hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
            HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&ssa::k_sweep_walk<P, F, false, 2, kWalkPrefetch>), hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
            HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&ssa::k_sweep_walk<P, F, false, 4, kWalkPrefetch>), hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
            HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&ssa::k_sweep_walk<P, F, false, 8, 8>), hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        }
        return D.walk_fits;
    }

    // out = M^{-1} in (ABMC order); t is the forward result; sb: the scaled
    // input's buffer (null: the handle's)
    void launch_precond(const P* in, P* out, P* tbuf, const SolverState* st, P* sb = nullptr)
    {
        if (scale_apply) {
            // S M^-1 S: the sweeps between two scalings
            P* sv = sb ? sb : sbuf.as<P>();
            launch_scale(in, sv);
            launch_sweeps(sv, out, tbuf, st);
            launch_scale(out, out);
            return;
        }
        launch_sweeps(in, out, tbuf, st);
    }

    void launch_scale(const P* in, P* out)
    {
        ssa::k_scale_rows<P><<<flat_grid(elements()), ssa::kBlock, 0, ctx->stream>>>(
            elements(), W, scale_d.as<double>(), in, out);
        HIP_CHECK(hipGetLastError());
    }

    // ---- vector plumbing: one vector (W == 1, batches included) or a block
    // of W right-hand sides in interleaved rows -------------------------------
    template <typename Fn>
    void with_width(Fn&& fn)
    {
        switch (W) {
        case 4: fn(std::integral_constant<int, 4>{}); break;
        case 8: fn(std::integral_constant<int, 8>{}); break;
        case 16: fn(std::integral_constant<int, 16>{}); break;
        default: throw SsaError(SSA_ERR_ARGUMENT, "unsupported block width");
        }
    }
    // element-wise kernels over count elements: up to 8192 workgroups
    static int flat_grid(size_t count)
    {
        return int(std::max<size_t>(1, std::min<size_t>((count + ssa::kBlock - 1) / ssa::kBlock, 8192)));
    }
    size_t elements() const { return size_t(n) * size_t(W); }

    // dst[i] = src[map[i]] for i < len (map -1: zero)
    void launch_gather(const index_t* map, const T* src, T* dst, index_t len)
    {
        ssa::k_gather<T><<<vec_grid(len), ssa::kBlock, 0, ctx->stream>>>(len, map, src, 0, dst, 0);
        HIP_CHECK(hipGetLastError());
    }

    // the caller's m vectors (tot values each, back to back) into device
    // order; block padding columns become zero
    void gather_in(const T* src, T* dst, int m)
    {
        if (W == 1) {
            launch_gather(rev.as<index_t>(), src, dst, n);
            return;
        }
        with_width([&](auto wc) {
            constexpr int WV = decltype(wc)::value;
            ssa::k_gather_blk<T, WV><<<flat_grid(elements()), ssa::kBlock, 0, ctx->stream>>>(
                n, rev.as<index_t>(), src, size_t(tot), m, dst);
        });
        HIP_CHECK(hipGetLastError());
    }

    // device order out to the caller's m vectors
    void scatter_out(const T* src, T* dst, int m)
    {
        if (W == 1) {
            launch_gather(perm.as<index_t>(), src, dst, tot);
            return;
        }
        with_width([&](auto wc) {
            constexpr int WV = decltype(wc)::value;
            ssa::k_scatter_blk<T, WV><<<flat_grid(size_t(tot) * size_t(m)), ssa::kBlock, 0, ctx->stream>>>(
                tot, perm.as<index_t>(), src, m, dst);
        });
        HIP_CHECK(hipGetLastError());
    }

    // preconditioner operands: CG applies M^{-1} to r (z out), COCR to q = Ap
    // (qt out); in mixed precision both go through the P copies lo_a, lo_b
    P* pre_in(bool cocr) { return kMixed ? lo_a.as<P>() : (cocr ? Ap.as<P>() : r.as<P>()); }
    P* pre_out(bool cocr) { return kMixed ? lo_b.as<P>() : (cocr ? qt.as<P>() : z.as<P>()); }
    P* pre_tmp() { return kMixed ? lo_t.as<P>() : t.as<P>(); }
    // the P copy a mixed-precision kernel also writes (none in FP64)
    P* low_copy() { return kMixed ? lo_a.as<P>() : nullptr; }
    // COCR's rt (P: the Krylov vector itself in FP64, the narrow copy in mixed)
    P* rt_vec() { return kMixed ? lo_r.as<P>() : z.as<P>(); }

    // dst = src converted (states that have stopped are left alone)
    template <typename S, typename D>
    void convert(const S* src, D* dst, const SolverState* st)
    {
        if (W == 1) {
            ssa::k_copy<S, D><<<vgrid(), ssa::kBlock, 0, ctx->stream>>>(n, src, dst, st, systems());
        } else {
            with_width([&](auto wc) {
                constexpr int WV = decltype(wc)::value;
                ssa::k_copy_blk<S, D, WV><<<flat_grid(elements()), ssa::kBlock, 0, ctx->stream>>>(
                    elements(), src, dst, st);
            });
        }
        HIP_CHECK(hipGetLastError());
    }

    // partials of conj?(a) . b per state (no fused control step)
    template <typename TA, typename TB>
    void launch_dot(const TA* a, const TB* b, int conj, T* part)
    {
        if (W == 1) {
            ssa::k_dot<T, TA, TB><<<vgrid(), ssa::kBlock, 0, ctx->stream>>>(n, a, b, conj, part, nullptr,
                                                                          ssa::Ctrl{}, systems());
        } else {
            with_width([&](auto wc) {
                constexpr int WV = decltype(wc)::value;
                ssa::k_dot_blk<T, TA, TB, WV><<<vgrid(), ssa::kBlock, 0, ctx->stream>>>(n, a, b, conj, part);
            });
        }
        HIP_CHECK(hipGetLastError());
    }

    // partials of |a|^2 (and |b|^2) per state
    void launch_norms(const T* a, const T* b, double* pa, double* pb)
    {
        if (W == 1) {
            ssa::k_norms<T><<<vgrid(), ssa::kBlock, 0, ctx->stream>>>(n, a, b, pa, pb, systems());
        } else {
            with_width([&](auto wc) {
                constexpr int WV = decltype(wc)::value;
                ssa::k_norms_blk<T, WV><<<vgrid(), ssa::kBlock, 0, ctx->stream>>>(n, a, b, pa, pb);
            });
        }
        HIP_CHECK(hipGetLastError());
    }

    // device rows [first, first + count) of a batch system's vector
    std::pair<size_t, size_t> rows_of(int q) const
    {
        if (count > 1) return {size_t(h_row0[q]), size_t(h_nrows[q])};
        return {0, size_t(n)};
    }
    // state q's part of src into dst (a batch system's rows, a block's column)
    void copy_rows(DevBuf& dst, const DevBuf& src, int q)
    {
        if (W > 1) {
            ssa::k_copy_column<T><<<vec_grid(n), ssa::kBlock, 0, ctx->stream>>>(n, W, q, src.as<T>(), dst.as<T>());
            HIP_CHECK(hipGetLastError());
            return;
        }
        const auto rr = rows_of(q);
        HIP_CHECK(hipMemcpyAsync(dst.as<T>() + rr.first, src.as<T>() + rr.first, rr.second * sizeof(T),
                                 hipMemcpyDeviceToDevice, ctx->stream));
    }

    // the caller's nrhs vectors (tot values each), in blocks of at most
    // kMaxRhs right-hand sides
    template <typename Fn>
    void in_blocks(int nrhs, Fn&& fn, int width = 0)
    {
        if (nrhs < 1) throw SsaError(SSA_ERR_ARGUMENT, "nrhs must be at least 1");
        if (nrhs > 1 && count > 1)
            throw SsaError(SSA_ERR_UNSUPPORTED, "several right-hand sides need a one-system handle");
        // right-hand sides per launch: the call's width, else the handle's
        const int blk = width > 0 ? width : rhs_block > 0 ? rhs_block : kMaxRhs;
        for (int r0 = 0; r0 < nrhs; r0 += blk) fn(r0, std::min(blk, nrhs - r0));
    }

    // ---- public operations -------------------------------------------------
    void apply(const void* xin, void* yout, bool on_device, int nrhs) override
    {
        in_blocks(nrhs, [&](int r0, int m) {
            const size_t off = size_t(r0) * size_t(tot);
            operate(static_cast<const T*>(xin) + off, static_cast<T*>(yout) + off, on_device, m, true);
        });
    }

    void spmv(const void* xin, void* yout, bool on_device, int nrhs) override
    {
        in_blocks(nrhs, [&](int r0, int m) {
            const size_t off = size_t(r0) * size_t(tot);
            operate(static_cast<const T*>(xin) + off, static_cast<T*>(yout) + off, on_device, m, false);
        });
    }

    // y = M^{-1} x (precond) or y = A x for m right-hand sides
    void operate(const T* xin, T* yout, bool on_device, int m, bool precond)
    {
        if (precond) no_split_apply("ssa_ic_apply");
        ensure_workspace(rhs_width(m));
        // the preconditioner's own apply: scaled on the level path only
        scale_apply = has_scale && !abmc_order;
        hipStream_t s = ctx->stream;
        const size_t vb = size_t(tot) * size_t(m) * sizeof(T);   // the caller's vectors
        const T* src = xin;
        if (!on_device) {
            HIP_CHECK(hipMemcpyAsync(tmp.p, xin, vb, hipMemcpyHostToDevice, s));
            src = tmp.as<T>();
        }
        gather_in(src, r.as<T>(), m);
        if (precond) {
            if (kMixed) convert(r.as<T>(), pre_in(false), nullptr);
            launch_precond(pre_in(false), pre_out(false), pre_tmp(), nullptr);
            if (kMixed) convert(pre_out(false), z.as<T>(), nullptr);
        } else {
            launch_spmv(r.as<T>(), z.as<T>(), 0, nullptr, nullptr);
        }
        T* dst = on_device ? yout : tmp2.as<T>();
        scatter_out(z.as<T>(), dst, m);
        if (!on_device) HIP_CHECK(hipMemcpyAsync(yout, tmp2.p, vb, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    }

    // bytes of scratch apply_noalloc carves: the stop states, the gathered
    // input (T), its preconditioner-precision copy (mixed), the forward
    // result, the output, the scaled input (level path with scaling), each
    // rounded up to 256 bytes
    size_t noalloc_bytes() const
    {
        auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
        const size_t vn = size_t(n);
        size_t b = up(sizeof(SolverState) * size_t(count)) + up(vn * sizeof(T)) + 2 * up(vn * sizeof(P));
        if (kMixed) b += up(vn * sizeof(P));
        if (has_scale && !abmc_order) b += up(vn * sizeof(P));
        return b;
    }

    // the SGS-MRTR split factor is no preconditioner of its own
    void no_split_apply(const char* who) const
    {
        if (factor_kind == sparsesolv_amd::kFactorSgsMrtr)
            throw SsaError(SSA_ERR_UNSUPPORTED,
                           std::string(who) + ": an SGS-MRTR handle has no preconditioner apply");
    }

    void apply_noalloc(const void* xin, void* yout) override
    {
        need_device();
        no_split_apply("ssa_ic_apply_noalloc");
        const size_t need = noalloc_bytes();
        if (!ctx->scratch || need > ctx->scratch_bytes)
            throw SsaError(SSA_ERR_ARGUMENT,
                           "ssa_ic_apply_noalloc: the system needs " + std::to_string(need) +
                               " bytes of scratch, more than the " + std::to_string(ctx->scratch_bytes) +
                               " the library holds per GPU (use ssa_ic_apply)");
        hipStream_t s = ctx->stream;
        char* q = static_cast<char*>(ctx->scratch);
        auto take = [&](size_t b) {
            char* r0 = q;
            q += (b + 255) & ~size_t(255);
            return r0;
        };
        const size_t vn = size_t(n);
        SolverState* st = reinterpret_cast<SolverState*>(take(sizeof(SolverState) * size_t(count)));
        T* a = reinterpret_cast<T*>(take(vn * sizeof(T)));
        P* pa = kMixed ? reinterpret_cast<P*>(take(vn * sizeof(P))) : reinterpret_cast<P*>(a);
        P* tb = reinterpret_cast<P*>(take(vn * sizeof(P)));
        P* ob = reinterpret_cast<P*>(take(vn * sizeof(P)));
        P* sb = (has_scale && !abmc_order) ? reinterpret_cast<P*>(take(vn * sizeof(P))) : nullptr;
        // one right-hand side, serial sweeps; the handle's own settings
        // come back on return
        struct Keep {
            IcImpl& h;
            int W;
            bool serial, scale;
            ~Keep()
            {
                h.W = W;
                h.serial_sweeps = serial;
                h.scale_apply = scale;
            }
        } keep{*this, W, serial_sweeps, scale_apply};
        W = 1;
        serial_sweeps = true;
        scale_apply = sb != nullptr;   // the preconditioner's own apply: scaled on the level path
        // the stop states start at zero (the array keeps the last call's contents)
        const size_t words = sizeof(SolverState) * size_t(count) / sizeof(unsigned);
        ssa::k_zero_words<<<1, ssa::kBlock, 0, s>>>(reinterpret_cast<unsigned*>(st), words);
        HIP_CHECK(hipGetLastError());
        launch_gather(rev.as<index_t>(), static_cast<const T*>(xin), a, n);
        if (kMixed) {
            ssa::k_copy<T, P><<<vgrid(), ssa::kBlock, 0, s>>>(n, a, pa, nullptr, systems());
            HIP_CHECK(hipGetLastError());
        }
        launch_precond(pa, ob, tb, st, sb);
        const T* res = reinterpret_cast<const T*>(ob);
        if (kMixed) {
            ssa::k_copy<P, T><<<vgrid(), ssa::kBlock, 0, s>>>(n, ob, a, nullptr, systems());
            HIP_CHECK(hipGetLastError());
            res = a;
        }
        launch_gather(perm.as<index_t>(), res, static_cast<T*>(yout), tot);
        HIP_CHECK(hipStreamSynchronize(s));
    }

    // Mixed-precision COCR carries rt = M^{-1} r by the recurrence
    // rt -= alpha qt, and qt comes from the float preconditioner: the error
    // of the recurrence grows until the iteration stalls (near 1e-7 of the
    // starting residual).  rt is therefore recomputed from the FP64
    // residual, rt = M^{-1} r (residual replacement), for the states whose
    // k_ctrl_replace decision says so (rst: status 0 for those); t = A rt
    // follows, so q = A p stays exact.  qt (lo_b) is consumed by then, lo_t
    // is scratch.
    void replace_rt(SolverState* st)
    {
        SolverState* rst = rstate.as<SolverState>();
        const int S = nstates();
        ssa::k_ctrl_replace<<<(S + 63) / 64, 64, 0, ctx->stream>>>(st, rst, S, kReplaceDrop, kReplaceGap);
        HIP_CHECK(hipGetLastError());
        st = rst;
        convert(r.as<T>(), lo_b.as<P>(), st);
        if (W == 1) {
            launch_precond(lo_b.as<P>(), rt_vec(), lo_t.as<P>(), st);
            return;
        }
        // a block's sweeps write every right-hand side of a row while any
        // runs: they go to scratch, and only the replacing ones are copied
        launch_precond(lo_b.as<P>(), lo_b.as<P>(), lo_t.as<P>(), st);
        convert(lo_b.as<P>(), rt_vec(), st);
    }
    // whether iteration k of a pass (counted from 0) replaces rt
    // whether iteration k of a pass (counted from 0) decides on a replacement
    static bool replace_at(int k) { return k % kGraphIterations == kGraphIterations - 1; }

    // one Krylov iteration (kernels only; no host synchronisation)
    void enqueue_iteration(int method, const SolveParams& prm, bool replace)
    {
        if (W > 1) {
            enqueue_block_iteration(method, prm, replace);
            return;
        }
        hipStream_t s = ctx->stream;
        SolverState* st = state.as<SolverState>();
        const int G = vgrid();
        const ssa::Systems y = systems();
        const bool pre = method == SSA_METHOD_ICCG || method == SSA_METHOD_ICCOCR ||
                         method == SSA_METHOD_NGS_ICCOCR;
        double* H = hist.as<double>();
        // control steps fused into the kernels that consume their result;
        // `spmv`: the partials come from the SpMV grid
        const int np_spmv = std::max(1, spmv_grid());
        auto ctl = [&](const void* part, int np, bool spmv) {
            return ssa::Ctrl{st, H, part, np, spmv ? 1 : 0, prm, 0, 0};
        };
        T* bx = prm.save_best ? best.as<T>() : nullptr;
        const bool c = prm.cocr != 0;
        if (!c) {
            // CGSolver::do_iterate body (cg_solver.hpp:85-140): the alpha step
            // runs in the update kernel, the convergence check in the r.z
            // kernel (after the preconditioner, which cannot change x or r),
            // beta and the best-x copy in the p update
            launch_spmv(p.as<T>(), Ap.as<T>(), prm.conjugate ? 2 : 1, part_a.as<T>(), st);
            traced(SSA_TRACE_UPDATE, s, vector_bytes(SSA_TRACE_UPDATE, false), [&] {
                ssa::k_cg_update<T, P><<<G, ssa::kBlock, 0, s>>>(n, x.as<T>(), r.as<T>(), p.as<T>(), Ap.as<T>(), part_b.as<double>(), st, ctl(part_a.p, np_spmv, true), low_copy(), y);
            });
            if (pre) launch_precond(pre_in(c), pre_out(c), pre_tmp(), st);
            else ssa::k_copy<T, T><<<G, ssa::kBlock, 0, s>>>(n, r.as<T>(), z.as<T>(), st, y);
            traced(SSA_TRACE_DOT, s, vector_bytes(SSA_TRACE_DOT, false), [&] {
                ssa::k_dot<T, T, P><<<G, ssa::kBlock, 0, s>>>(n, r.as<T>(), pre_out(c), prm.conjugate, part_c.as<T>(), st, ctl(part_b.p, G, false), y);
            });
            traced(SSA_TRACE_PQ, s, vector_bytes(SSA_TRACE_PQ, false), [&] {
                ssa::k_xpby<T, P><<<G, ssa::kBlock, 0, s>>>(n, pre_out(c), p.as<T>(), x.as<T>(), bx, st, ctl(part_c.p, G, false), y);
            });
        } else {
            // COCRSolver::do_iterate body (cocr_solver.hpp:96-162); z holds rt,
            // Ap holds q.  alpha in the update kernel, the check in the
            // t = A rt kernel, beta and the best-x copy in the p, q update
            if (pre) launch_precond(pre_in(c), pre_out(c), pre_tmp(), st);
            else ssa::k_copy<T, T><<<G, ssa::kBlock, 0, s>>>(n, Ap.as<T>(), qt.as<T>(), st, y);
            traced(SSA_TRACE_DOT, s, vector_bytes(SSA_TRACE_DOT, true), [&] {
                if (kMixed)   // (q, qt) with the narrow operand second
                    ssa::k_dot<T, T, P><<<G, ssa::kBlock, 0, s>>>(n, Ap.as<T>(), pre_out(c), 0, part_a.as<T>(), st, ssa::Ctrl{}, y);
                else
                    ssa::k_dot<T, P, T><<<G, ssa::kBlock, 0, s>>>(n, pre_out(c), Ap.as<T>(), 0, part_a.as<T>(), st, ssa::Ctrl{}, y);
            });
            traced(SSA_TRACE_UPDATE, s, vector_bytes(SSA_TRACE_UPDATE, true), [&] {
                ssa::k_cocr_update<T, P, P><<<G, ssa::kBlock, 0, s>>>(n, x.as<T>(), r.as<T>(), rt_vec(), p.as<T>(), Ap.as<T>(), pre_out(c), part_b.as<double>(), st, ctl(part_a.p, G, false), y, bx);
            });
            if (kMixed && pre && replace) replace_rt(st);
            launch_spmv_low(rt_vec(), t.as<T>(), 1, part_c.as<T>(), st, ctl(part_b.p, G, false));
            traced(SSA_TRACE_PQ, s, vector_bytes(SSA_TRACE_PQ, true), [&] {
                ssa::k_cocr_pq<T, P, P><<<G, ssa::kBlock, 0, s>>>(n, rt_vec(), t.as<T>(), p.as<T>(), Ap.as<T>(), st, ctl(part_c.p, np_spmv, true), low_copy(), y);
            });
        }
        HIP_CHECK(hipGetLastError());
    }

    // The same iteration for a block of W right-hand sides: the same kernels'
    // arithmetic per right-hand side, with each control step in its own
    // W-workgroup kernel between producer and consumer (same state updates in
    // the same order: alpha, check, beta).
    void enqueue_block_iteration(int method, const SolveParams& prm, bool replace)
    {
        hipStream_t s = ctx->stream;
        SolverState* st = state.as<SolverState>();
        const int G = vgrid();
        const bool pre = method == SSA_METHOD_ICCG || method == SSA_METHOD_ICCOCR ||
                         method == SSA_METHOD_NGS_ICCOCR;
        double* H = hist.as<double>();
        const int hs = hist_stride;
        const int np_spmv = std::max(1, spmv_grid());
        T* bx = prm.save_best ? best.as<T>() : nullptr;
        const bool c = prm.cocr != 0;
        const size_t ne = elements();
        const int FG = flat_grid(ne);
        constexpr int CT = ssa::kCtrlThreads;
        with_width([&](auto wc) {
            constexpr int WV = decltype(wc)::value;
            if (!c) {
                launch_spmv(p.as<T>(), Ap.as<T>(), prm.conjugate ? 2 : 1, part_a.as<T>(), st);
                ssa::k_ctrl_alpha_blk<T><<<WV, CT, 0, s>>>(part_a.as<T>(), np_spmv, st, prm, H, hs);
                traced(SSA_TRACE_UPDATE, s, vector_bytes(SSA_TRACE_UPDATE, false), [&] {
                    ssa::k_cg_update_blk<T, P, WV><<<G, ssa::kBlock, 0, s>>>(n, x.as<T>(), r.as<T>(), p.as<T>(), Ap.as<T>(), part_b.as<double>(), st, low_copy());
                });
                ssa::k_ctrl_check_blk<<<WV, CT, 0, s>>>(part_b.as<double>(), G, st, prm, H, hs);
                if (pre) launch_precond(pre_in(c), pre_out(c), pre_tmp(), st);
                else ssa::k_copy_blk<T, T, WV><<<FG, ssa::kBlock, 0, s>>>(ne, r.as<T>(), z.as<T>(), st);
                traced(SSA_TRACE_DOT, s, vector_bytes(SSA_TRACE_DOT, false), [&] {
                    ssa::k_dot_blk<T, T, P, WV><<<G, ssa::kBlock, 0, s>>>(n, r.as<T>(), pre_out(c), prm.conjugate, part_c.as<T>());
                });
                ssa::k_ctrl_beta_blk<T><<<WV, CT, 0, s>>>(part_c.as<T>(), G, st, prm);
                traced(SSA_TRACE_PQ, s, vector_bytes(SSA_TRACE_PQ, false), [&] {
                    ssa::k_xpby_blk<T, P, WV><<<FG, ssa::kBlock, 0, s>>>(ne, pre_out(c), p.as<T>(), x.as<T>(), bx, st);
                });
            } else {
                if (pre) launch_precond(pre_in(c), pre_out(c), pre_tmp(), st);
                else ssa::k_copy_blk<T, T, WV><<<FG, ssa::kBlock, 0, s>>>(ne, Ap.as<T>(), qt.as<T>(), st);
                traced(SSA_TRACE_DOT, s, vector_bytes(SSA_TRACE_DOT, true), [&] {
                    if (kMixed)   // (q, qt) with the narrow operand second
                        ssa::k_dot_blk<T, T, P, WV><<<G, ssa::kBlock, 0, s>>>(n, Ap.as<T>(), pre_out(c), 0, part_a.as<T>());
                    else
                        ssa::k_dot_blk<T, P, T, WV><<<G, ssa::kBlock, 0, s>>>(n, pre_out(c), Ap.as<T>(), 0, part_a.as<T>());
                });
                ssa::k_ctrl_alpha_blk<T><<<WV, CT, 0, s>>>(part_a.as<T>(), G, st, prm, H, hs);
                traced(SSA_TRACE_UPDATE, s, vector_bytes(SSA_TRACE_UPDATE, true), [&] {
                    ssa::k_cocr_update_blk<T, P, WV, P><<<G, ssa::kBlock, 0, s>>>(n, x.as<T>(), r.as<T>(), rt_vec(), p.as<T>(), Ap.as<T>(), pre_out(c), part_b.as<double>(), st, bx);
                });
                ssa::k_ctrl_check_blk<<<WV, CT, 0, s>>>(part_b.as<double>(), G, st, prm, H, hs);
                if (kMixed && pre && replace) replace_rt(st);
                launch_spmv_low(rt_vec(), t.as<T>(), 1, part_c.as<T>(), st);
                ssa::k_ctrl_beta_blk<T><<<WV, CT, 0, s>>>(part_c.as<T>(), np_spmv, st, prm);
                traced(SSA_TRACE_PQ, s, vector_bytes(SSA_TRACE_PQ, true), [&] {
                    ssa::k_cocr_q_blk<T, P, WV><<<FG, ssa::kBlock, 0, s>>>(ne, t.as<T>(), Ap.as<T>(), st, low_copy());
                });
            }
        });
        HIP_CHECK(hipGetLastError());
    }

    // ||v|| of every state in FP64 (partials summed on the host in index
    // order, as a one-system solve of that system sums them)
    std::vector<double> state_norms(const T* v)
    {
        hipStream_t s = ctx->stream;
        launch_norms(v, nullptr, part_d.as<double>(), nullptr);
        const int G = vgrid();
        std::vector<double> h(size_t(G) * size_t(W), 0.0);
        HIP_CHECK(hipMemcpyAsync(h.data(), part_d.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        const int S = nstates();
        std::vector<double> out(size_t(S), 0.0);
        for (int q = 0; q < S; ++q) {
            const int a = count > 1 ? h_vec_g0[q] : q * G, b = count > 1 ? h_vec_g0[q + 1] : (q + 1) * G;
            double sum = 0.0;
            for (int i = a; i < b; ++i) sum += h[size_t(i)];
            out[size_t(q)] = std::sqrt(sum);
        }
        return out;
    }

    // ||b - A x|| of every state in FP64 (uses Ap and r)
    std::vector<double> true_residuals()
    {
        const index_t rows = index_t(elements());
        launch_spmv(x.as<T>(), Ap.as<T>(), 0, nullptr, nullptr);
        ssa::k_residual<T><<<vec_grid(rows), ssa::kBlock, 0, ctx->stream>>>(rows, bvec.as<T>(), Ap.as<T>(), r.as<T>());
        HIP_CHECK(hipGetLastError());
        return state_norms(r.as<T>());
    }

    // prepare_iteration and the iteration loop from the current x, with the
    // per-state settings in h_pass; leaves the final states in h_state
    void run_pass(int method, const SolveParams& prm, bool pre, const ssa_config& cfg)
    {
        hipStream_t s = ctx->stream;
        const int G = vgrid();
        const int S = nstates();
        const ssa::Systems y = systems();
        SolverState* st = state.as<SolverState>();
        double* H = hist.as<double>();
        HIP_CHECK(hipMemsetAsync(state.p, 0, sizeof(SolverState) * S, s));
        HIP_CHECK(hipMemcpyAsync(pass.p, h_pass, sizeof(ssa::PassParams) * S, hipMemcpyHostToDevice, s));
        int maxit = 0;
        for (int q = 0; q < S; ++q)
            if (h_pass[q].active) maxit = std::max(maxit, h_pass[q].max_iterations);

        // prepare_iteration: r = b - A x, norms, normaliser (one control
        // workgroup per state)
        const index_t rows = index_t(elements());
        launch_spmv(x.as<T>(), Ap.as<T>(), 0, nullptr, nullptr);
        ssa::k_residual<T><<<vec_grid(rows), ssa::kBlock, 0, s>>>(rows, bvec.as<T>(), Ap.as<T>(), r.as<T>());
        launch_norms(bvec.as<T>(), r.as<T>(), part_b.as<double>(), part_d.as<double>());
        ssa::k_ctrl_init<<<S, ssa::kBlock, 0, s>>>(part_b.as<double>(), part_d.as<double>(), G, st, prm, H, y,
                                                   pass.as<ssa::PassParams>());
        HIP_CHECK(hipGetLastError());
        const bool c = prm.cocr != 0;
        if (!c) {
            // rz_old = (r, M^{-1} r); p = z (cg_solver.hpp:73-77)
            if (pre) {
                if (kMixed) convert(r.as<T>(), pre_in(c), nullptr);
                launch_precond(pre_in(c), pre_out(c), pre_tmp(), nullptr);
            } else {
                convert(r.as<T>(), z.as<T>(), nullptr);
            }
            launch_dot(r.as<T>(), pre_out(c), prm.conjugate, part_c.as<T>());
            ssa::k_ctrl_store<T><<<S, ssa::kBlock, 0, s>>>(part_c.as<T>(), G, st, 0, y, 0);
            convert(pre_out(c), p.as<T>(), nullptr);
        } else {
            // rt = M^{-1} r; p = rt; q = A p; rho = rt^T q (cocr_solver.hpp:78-88)
            if (pre) {
                if (kMixed) convert(r.as<T>(), pre_in(c), nullptr);
                launch_precond(kMixed ? pre_in(c) : r.as<P>(), rt_vec(), pre_tmp(), nullptr);
            } else {
                convert(r.as<T>(), z.as<T>(), nullptr);
            }
            convert(rt_vec(), p.as<T>(), nullptr);
            launch_spmv(p.as<T>(), Ap.as<T>(), 1, part_c.as<T>(), nullptr);
            ssa::k_ctrl_store<T><<<S, ssa::kBlock, 0, s>>>(part_c.as<T>(), std::max(1, spmv_grid()), st, 0, y, 1);
            if (kMixed) convert(Ap.as<T>(), pre_in(c), nullptr);
        }
        HIP_CHECK(hipGetLastError());

        iterate(maxit, cfg, [&](int k) { enqueue_iteration(method, prm, replace_at(k)); });
    }

    // The iteration loop: replay a captured group of iterations, poll the
    // stop status between groups.  The control kernels stop each state at
    // its iteration budget themselves and every kernel returns at once after
    // a stop, so iterations enqueued past the stop cost only launches.
    // enqueue(k) enqueues iteration k of the pass (counted from 0); leaves
    // the final states in h_state.
    template <typename Fn>
    void iterate(int maxit, const ssa_config& cfg, Fn&& enqueue)
    {
        hipStream_t s = ctx->stream;
        const int S = nstates();
        SolverState* st = state.as<SolverState>();
        // timed launches (ssa_ic_trace) need the eager loop
        const bool graph_on = cfg.use_graph && trace_cls < 0;
        const int unroll = graph_on ? std::min(kGraphIterations, maxit + 1) : 1;
        int poll = std::max(cfg.poll_interval > 0 ? cfg.poll_interval : 8, unroll);
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        if (graph_on) {
            HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int u = 0; u < unroll; ++u) enqueue(u);
            HIP_CHECK(hipStreamEndCapture(s, &graph));
            HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        }
        int done = 0, enqueued = 0;
        // runs while any state runs (a stopped state's kernels return at once)
        auto running = [&] {
            HIP_CHECK(hipMemcpyAsync(h_state, st, sizeof(SolverState) * S, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            trace_harvest();
            for (int q = 0; q < S; ++q)
                if (h_state[q].status == ssa::kRunning) return true;
            return false;
        };
        bool more = running();
        struct Disarm {
            bool& v;
            ~Disarm() { v = false; }
        } disarm{trace_armed};
        while (more && done <= maxit) {
            const int k = std::min(poll, maxit + 1 - done);
            const int groups = (k + unroll - 1) / unroll;
            trace_armed = trace_cls >= 0;
            for (int i = 0; i < groups; ++i) {
                if (exec) HIP_CHECK(hipGraphLaunch(exec, s));
                else enqueue(enqueued++);
            }
            trace_armed = false;
            done += groups * unroll;
            more = running();
            if (poll < 64) poll *= 2;
        }
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }

    // solve_sgsmrtr (sparsesolv.hpp:123-137): SGSMRTRSolver's
    // prepare_iteration and do_iterate (solvers/sgs_mrtr_solver.hpp:82-260)
    // in device order.  bvec holds b and x the initial guess on entry; on
    // return x holds x2 = D^-1 x (the iterate in the scaled space: the
    // caller takes the best one when the solve did not converge, then
    // x = D x2) and h_state the final state.  Vectors: r = rd, z = u,
    // qt = the forward input (r = D (b - A x) first, then rd - u), Ap = A x
    // first, then w and Ard, t = y, tmp = rd / L_ii, p = p, bvec = b2.
    void run_mrtr(const SolveParams& prm, const ssa_config& cfg)
    {
        if constexpr (kMixed) {
            throw SsaError(SSA_ERR_UNSUPPORTED, "SGS-MRTR runs in FP64");
        } else {
            hipStream_t s = ctx->stream;
            const int G = vgrid();
            SolverState* st = state.as<SolverState>();
            HIP_CHECK(hipMemsetAsync(state.p, 0, sizeof(SolverState), s));
            HIP_CHECK(hipMemsetAsync(p.p, 0, size_t(n) * sizeof(T), s));
            launch_spmv(x.as<T>(), Ap.as<T>(), 0, nullptr, nullptr);
            ssa::k_mrtr_prep<T><<<G, ssa::kBlock, 0, s>>>(n, mr_d.as<T>(), mr_id.as<T>(), bvec.as<T>(), Ap.as<T>(), x.as<T>(), qt.as<T>(), part_a.as<double>(), part_b.as<double>());
            ssa::k_mrtr_lnorm<T><<<G, ssa::kBlock, 0, s>>>(n, mr_rp.as<index_t>(), mr_ci.as<index_t>(), mr_v.as<T>(), bvec.as<T>(), part_c.as<double>());
            HIP_CHECK(hipGetLastError());
            launch_sweep<true>(fwd, qt.as<T>(), r.as<T>(), st);   // rd = L^-1 r
            ssa::k_mrtr_start<T><<<G, ssa::kBlock, 0, s>>>(n, r.as<T>(), invd.as<T>(), t.as<T>(), tmp.as<T>(), part_d.as<double>());
            ssa::k_mrtr_init<<<1, ssa::kCtrlThreads, 0, s>>>(part_a.as<double>(), part_b.as<double>(), part_c.as<double>(), part_d.as<double>(), G, st, prm, hist.as<double>());
            HIP_CHECK(hipGetLastError());
            iterate(prm.max_iterations, cfg, [&](int) { enqueue_mrtr_iteration(prm); });
        }
    }

    // one SGS-MRTR iteration (sgs_mrtr_solver.hpp:166-240)
    void enqueue_mrtr_iteration(const SolveParams& prm)
    {
        if constexpr (!kMixed) {
            hipStream_t s = ctx->stream;
            SolverState* st = state.as<SolverState>();
            const int G = vgrid();
            constexpr int CT = ssa::kCtrlThreads;
            launch_sweep<false>(bwd, tmp.as<T>(), z.as<T>(), st);   // u = L^-T rd
            ssa::k_mrtr_sub<T><<<G, ssa::kBlock, 0, s>>>(n, r.as<T>(), z.as<T>(), qt.as<T>(), st);
            launch_sweep<true>(fwd, qt.as<T>(), Ap.as<T>(), st);    // w = L^-1 (rd - u)
            ssa::k_mrtr_dots<T><<<G, ssa::kBlock, 0, s>>>(n, Ap.as<T>(), z.as<T>(), r.as<T>(), t.as<T>(), prm.conjugate, part_a.as<T>(), part_b.as<T>(), part_c.as<T>(), st);
            ssa::k_mrtr_coef<T><<<1, CT, 0, s>>>(part_a.as<T>(), part_b.as<T>(), part_c.as<T>(), G, st);
            ssa::k_mrtr_update<T><<<G, ssa::kBlock, 0, s>>>(n, z.as<T>(), Ap.as<T>(), p.as<T>(), x.as<T>(), t.as<T>(), r.as<T>(), tmp.as<T>(), invd.as<T>(), part_d.as<double>(), st);
            ssa::k_mrtr_check<<<1, CT, 0, s>>>(part_d.as<double>(), G, st, prm, hist.as<double>());
            if (prm.save_best) ssa::k_mrtr_best<T><<<G, ssa::kBlock, 0, s>>>(n, x.as<T>(), best.as<T>(), st);
            HIP_CHECK(hipGetLastError());
        }
    }

    void solve(int method, const void* bin, void* xio, bool on_device, const ssa_config& cfg,
               ssa_result* results, double* history, int nrhs) override
    {
        in_blocks(nrhs, [&](int r0, int m) {
            const size_t off = size_t(r0) * size_t(tot);
            solve_block(method, static_cast<const T*>(bin) + off, static_cast<T*>(xio) + off, on_device,
                        cfg, results + r0, nrhs == 1 ? history : nullptr, m);
        }, cfg.rhs_block);
    }

    // one block of m right-hand sides (m = 1 for a batch: its systems)
    void solve_block(int method, const T* bin, T* xio, bool on_device, const ssa_config& cfg,
                     ssa_result* results, double* history, int m)
    {
        ensure_workspace(rhs_width(m));
        // scaled apply: the level path, and ICCG in the reordered space
        // (Path 1, sparsesolv.hpp:58-76) or in the RCM space (Path 2,
        // :78-96, whose adapter scales, ic_preconditioner.hpp:327-340); the
        // others use the ABMC apply
        scale_apply = has_scale &&
                      (!abmc_order || (method == SSA_METHOD_ICCG &&
                                       (cfg.abmc_reorder_spmv || abmc_rcm)));
        hipStream_t s = ctx->stream;
        const int S = nstates();
        const int nres = count > 1 ? count : m;      // results the caller receives
        const size_t vb = size_t(tot) * size_t(m) * sizeof(T);   // the caller's vectors
        const size_t db = size_t(n) * size_t(W) * sizeof(T);     // device rows
        SolveParams prm{};
        prm.tolerance = cfg.tolerance;
        prm.abs_tolerance = cfg.abs_tolerance;
        prm.custom_norm = cfg.custom_norm;
        prm.divergence_threshold = cfg.divergence_threshold;
        prm.max_iterations = cfg.max_iterations;
        prm.norm_type = cfg.norm_type;
        prm.save_best = cfg.save_best_result ? 1 : 0;
        prm.save_history = cfg.save_residual_history ? 1 : 0;
        prm.divergence_check = cfg.divergence_check ? 1 : 0;
        prm.divergence_count = cfg.divergence_count;
        prm.conjugate = cfg.conjugate ? 1 : 0;
        prm.cocr = (method == SSA_METHOD_ICCOCR || method == SSA_METHOD_COCR ||
                    method == SSA_METHOD_NGS_ICCOCR) ? 1 : 0;
        prm.ngs = method == SSA_METHOD_NGS_ICCOCR ? 1 : 0;
        if (prm.ngs) {
            prm.save_best = 0;
            prm.divergence_check = 0;
        }
        const bool pre = method == SSA_METHOD_ICCG || method == SSA_METHOD_ICCOCR ||
                         method == SSA_METHOD_NGS_ICCOCR;
        hist_stride = std::max(cfg.max_iterations, 0) + 4;
        hist.reserve(sizeof(double) * size_t(hist_stride) * size_t(S));
        double* H = hist.as<double>();

        if (S > 1 && prm.ngs)
            throw SsaError(SSA_ERR_UNSUPPORTED, "COCRSolverNGS runs one system and one right-hand side");
        const bool mrtr = method == SSA_METHOD_SGSMRTR;
        if (mrtr != (factor_kind == sparsesolv_amd::kFactorSgsMrtr))
            throw SsaError(SSA_ERR_ARGUMENT,
                           mrtr ? "SSA_METHOD_SGSMRTR needs a handle created with preconditioner = "
                                  "SSA_PRECOND_SGSMRTR"
                                : "an SSA_PRECOND_SGSMRTR handle solves with SSA_METHOD_SGSMRTR only");
        if (mrtr && S > 1)
            throw SsaError(SSA_ERR_UNSUPPORTED, "SGS-MRTR runs one right-hand side");
        if (kMixed && !pre)
            throw SsaError(SSA_ERR_UNSUPPORTED,
                           "ssa_solve: mixed precision needs the IC preconditioner");
        const auto t0 = std::chrono::steady_clock::now();
        // inputs into ABMC order: b -> bvec, x0 -> x
        const T* bsrc = bin;
        const T* xsrc = xio;
        if (!on_device) {
            HIP_CHECK(hipMemcpyAsync(tmp.p, bin, vb, hipMemcpyHostToDevice, s));
            HIP_CHECK(hipMemcpyAsync(tmp2.p, xio, vb, hipMemcpyHostToDevice, s));
            bsrc = tmp.as<T>();
            xsrc = tmp2.as<T>();
        }
        gather_in(bsrc, bvec.as<T>(), m);
        if (prm.ngs) HIP_CHECK(hipMemsetAsync(x.p, 0, db, s));
        else gather_in(xsrc, x.as<T>(), m);
        HIP_CHECK(hipMemsetAsync(best.p, 0, db, s));

        // per state: whether it is solved at all (block padding is not),
        // iterations, the FP64 true residual (mixed) and the final state
        std::vector<char> active(size_t(S), 1);
        for (int q = nres; q < S; ++q) active[size_t(q)] = 0;
        std::vector<int> iters(size_t(S), 0);
        std::vector<double> true_rel(size_t(S), -1.0);
        std::vector<SolverState> fin(size_t(S), SolverState{});
        if (!kMixed || prm.ngs) {
            for (int q = 0; q < S; ++q)
                h_pass[q] = ssa::PassParams{cfg.tolerance, cfg.max_iterations, active[size_t(q)]};
            if (mrtr) run_mrtr(prm, cfg);
            else run_pass(method, prm, pre, cfg);
            for (int q = 0; q < S; ++q) {
                fin[size_t(q)] = h_state[q];
                iters[size_t(q)] = h_state[q].result_iters;
            }
        } else {
            // Mixed precision, iterative refinement, per state: the residual
            // is computed in FP64 from x, each pass runs the preconditioned
            // Krylov method from x until its residual falls below the
            // tolerance (COCR with residual replacement, replace_rt), and
            // the next pass restarts from the new x.  A pass that ends
            // unconverged leaves its best iterate (when save_best_result); a
            // pass that does not lower the true residual is undone, so x and
            // the reported residual always match.  A state leaves the loop
            // when it converges, runs out of iterations, or stops improving.
            const std::vector<double> bnorm = state_norms(bvec.as<T>());
            std::vector<double> scale(static_cast<size_t>(S)), rel(static_cast<size_t>(S));
            std::vector<double> res0 = true_residuals();
            for (int q = 0; q < S; ++q) {
                scale[size_t(q)] = bnorm[size_t(q)] > 0.0 ? bnorm[size_t(q)] : 1.0;
                rel[size_t(q)] = res0[size_t(q)] / scale[size_t(q)];
            }
            std::vector<char> done(size_t(S), 0), run(size_t(S), 0);
            for (int pass = 0; pass < kRefinePasses; ++pass) {
                bool any = false;
                for (int q = 0; q < S; ++q) {
                    const size_t u = size_t(q);
                    run[u] = active[u] && !done[u] && !(rel[u] < cfg.tolerance) &&
                             iters[u] < cfg.max_iterations;
                    any = any || run[u];
                    h_pass[q] = ssa::PassParams{cfg.tolerance, cfg.max_iterations - iters[u], run[u] ? 1 : 0};
                }
                if (!any) break;
                HIP_CHECK(hipMemcpyAsync(tmp.p, x.p, db, hipMemcpyDeviceToDevice, s));
                run_pass(method, prm, pre, cfg);
                for (int q = 0; q < S; ++q) {
                    const size_t u = size_t(q);
                    if (!run[u]) continue;
                    fin[u] = h_state[q];
                    iters[u] += h_state[q].result_iters;
                    // the best iterate: x itself when the last check improved
                    // (the copy into best follows an improving check)
                    if (!h_state[q].converged && prm.save_best && !h_state[q].improved)
                        copy_rows(x, best, q);
                }
                const std::vector<double> res = true_residuals();
                for (int q = 0; q < S; ++q) {
                    const size_t u = size_t(q);
                    if (!run[u]) continue;
                    const double next = res[u] / scale[u];
                    if (!(next < rel[u])) {
                        copy_rows(x, tmp, q);
                        done[u] = 1;
                        continue;
                    }
                    if (!(next < 0.5 * rel[u])) done[u] = 1;
                    rel[u] = next;
                }
            }
            for (int q = 0; q < S; ++q) {
                const size_t u = size_t(q);
                SolverState& Sq = fin[u];
                true_rel[u] = rel[u];
                Sq.converged = rel[u] < cfg.tolerance;
                if (Sq.converged) Sq.status = ssa::kConverged;
                else if (Sq.status == ssa::kRunning) Sq.status = ssa::kMaxIter;
            }
        }

        // build_result (iterative_solver.hpp:262-279), per state
        for (int q = 0; q < nres; ++q) {
            const size_t u = size_t(q);
            const SolverState& Sq = fin[u];
            ssa_result& res = results[q];
            res.converged = Sq.converged;
            res.iterations = iters[u];
            res.status = Sq.status;
            res.final_residual = true_rel[u] >= 0.0 ? true_rel[u] : Sq.final_norm / Sq.normalizer;
            res.history_len = prm.save_history && S == 1 ? Sq.hist_len : 0;
            res.actual_shift = pre ? actual_shift : 0.0;
            if (true_rel[u] < 0.0 && !Sq.converged && prm.save_best) {
                // x is the best iterate when the last check improved; best
                // holds it otherwise (copied after its check)
                if (!Sq.improved) copy_rows(x, best, q);
                res.final_residual = Sq.best_residual;
            }
        }
        if (mrtr) {
            ssa::k_mrtr_unscale<T><<<vgrid(), ssa::kBlock, 0, s>>>(n, mr_d.as<T>(), x.as<T>());
            HIP_CHECK(hipGetLastError());
        }
        T* xdst = on_device ? xio : tmp2.as<T>();
        scatter_out(x.as<T>(), xdst, m);
        if (!on_device) HIP_CHECK(hipMemcpyAsync(xio, tmp2.p, vb, hipMemcpyDeviceToHost, s));
        if (history && S == 1 && results[0].history_len > 0)
            HIP_CHECK(hipMemcpyAsync(history, H, sizeof(double) * results[0].history_len, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (int q = 0; q < nres; ++q) results[q].solve_seconds = secs;
    }

    // Each colour launched on its own (no tail group, no oversized-block
    // work) is run `reps` times back to back between two events, after the
    // whole sweep has run once; avg_us is the mean over those colours of one
    // launch's share, bytes the mean of their algorithmic bytes.  Zero when
    // no colour qualifies.
    void profile_colour_kernels(int reps, double& avg_us, double& bytes)
    {
        hipStream_t s = ctx->stream;
        hipEvent_t e0, e1;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        double total_us = 0.0, total_bytes = 0.0;
        int count = 0;
        for (int dir = 0; dir < 2; ++dir) {
            DevSweep& D = dir == 0 ? fwd : bwd;
            P* in = dir == 0 ? r.as<P>() : t.as<P>();
            P* out = dir == 0 ? t.as<P>() : z.as<P>();
            if (dir == 0) launch_sweep<true>(D, in, out, nullptr);
            else launch_sweep<false>(D, in, out, nullptr);
            for (index_t ph = 0; ph < D.nphases; ++ph) {
                if (in_tail_group(D, ph) || D.phase_big[ph + 1] > D.phase_big[ph] ||
                    D.phase_slices[ph + 1] == D.phase_slices[ph])
                    continue;
                HIP_CHECK(hipEventRecord(e0, s));
                for (int k = 0; k < reps; ++k) {
                    if (dir == 0) launch_phase<true>(D, ph, in, out, nullptr);
                    else launch_phase<false>(D, ph, in, out, nullptr);
                }
                HIP_CHECK(hipEventRecord(e1, s));
                HIP_CHECK(hipEventSynchronize(e1));
                float ms = 0.f;
                HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                total_us += 1000.0 * double(ms) / reps;
                total_bytes += colour_bytes(D, ph);
                ++count;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        avg_us = count ? total_us / count : 0.0;
        bytes = count ? total_bytes / count : 0.0;
    }

    void profile(int what, int reps, int nrhs, double& avg_us, double& bytes) override
    {
        if (nrhs < 1 || nrhs > kMaxRhs)
            throw SsaError(SSA_ERR_ARGUMENT, "ssa_ic_profile: nrhs must be 1.." + std::to_string(kMaxRhs));
        ensure_workspace(rhs_width(nrhs));
        scale_apply = false;   // the sweeps themselves are timed
        hipStream_t s = ctx->stream;
        // operands of all W right-hand sides (a block pads nrhs up to W)
        const size_t rows = size_t(n) * size_t(W);
        const size_t vb = rows * sizeof(T);
        const size_t pb = rows * sizeof(P);
        // deterministic non-zero operands
        std::vector<HT> h(rows);
        for (size_t i = 0; i < rows; ++i) h[i] = HT(1.0 / (1.0 + double(i % 97)));
        HIP_CHECK(hipMemcpyAsync(p.p, h.data(), vb, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(r.p, h.data(), vb, hipMemcpyHostToDevice, s));
        auto run = [&] {
            switch (what) {
            case SSA_PROFILE_SPMV: launch_spmv(p.as<T>(), Ap.as<T>(), 1, part_a.as<T>(), nullptr); break;
            case SSA_PROFILE_FORWARD:
                launch_sweep<true>(fwd, r.as<P>(), t.as<P>(), nullptr);
                break;
            case SSA_PROFILE_BACKWARD:
                launch_sweep<false>(bwd, t.as<P>(), z.as<P>(), nullptr);
                break;
            default: launch_precond(r.as<P>(), z.as<P>(), t.as<P>(), nullptr); break;
            }
        };
        run();   // warm-up
        if (what == SSA_PROFILE_COLOUR_KERNELS) {
            profile_colour_kernels(std::max(reps, 1), avg_us, bytes);
            return;
        }
        hipEvent_t e0, e1;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        HIP_CHECK(hipEventRecord(e0, s));
        for (int k = 0; k < reps; ++k) run();
        HIP_CHECK(hipEventRecord(e1, s));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        avg_us = 1000.0 * double(ms) / std::max(reps, 1);
        // algorithmic bytes: matrix entries (value + int32 column) once, plus
        // row pointers, plus every vector operand once
        const double rowptr = 4.0 * (double(n) + 1.0);
        const double lbytes = double(nnz_l) * (sizeof(F) + 4.0) + rowptr;
        const double abytes = double(nnz) * (sizeof(T) + 4.0) + rowptr;
        switch (what) {
        case SSA_PROFILE_SPMV: bytes = abytes + 2.0 * double(vb); break;            // read x, write y
        case SSA_PROFILE_FORWARD: bytes = lbytes + 2.0 * double(pb) + double(n) * sizeof(F); break;
        case SSA_PROFILE_BACKWARD: bytes = lbytes + 2.0 * double(pb) + double(n) * sizeof(F); break;
        default: bytes = 2.0 * lbytes + 4.0 * double(pb) + 2.0 * double(n) * sizeof(F); break;
        }
    }

    // the operator of the solves and the SpMV: another matrix of the same
    // order, in the device order (the factor and its layouts stay)
    void set_operator(int nn, const int32_t* rp, const int32_t* ci, const void* vals) override
    {
        need_device();
        if (count != 1)
            throw SsaError(SSA_ERR_UNSUPPORTED, "ssa_ic_set_operator: not available for a batched handle");
        if (nn != n)
            throw SsaError(SSA_ERR_ARGUMENT, "ssa_ic_set_operator: the operator must have the handle's order");
        const auto B = sparsesolv_amd::permute_csr<HT>(nn, rp, ci, static_cast<const HT*>(vals), h_ordering);
        sparsesolv_amd::SpmvLayout<HT> L;
        L.build(B, h_ordering);
        hipStream_t s = ctx->stream;
        HIP_CHECK(hipStreamSynchronize(s));   // queued work may still read the old layout
        A_off.upload(L.slice_off, s);
        A_val.upload(L.val, s);
        A_col.upload(L.col, s);
        A_row.upload(L.row, s);
        HIP_CHECK(hipStreamSynchronize(s));   // the host arrays end with this call
        A_nslices = L.nslices;
        spmv_bytes = L.val.size() * sizeof(HT) + L.col.size() * 4;
        nnz = nn > 0 ? rp[nn] : 0;
    }

    void info(ssa_ic_info& o) const override
    {
        o.n = tot;   // batches: all systems' rows (colours: the most of any system)
        o.nnz = nnz;
        o.nnz_l = nnz_l;
        o.num_colors = nc;
        o.num_blocks = nb;
        o.rows_per_lane = fwd.BR;
        o.big_blocks = fwd.phase_big.empty() ? 0 : fwd.phase_big.back();
        o.max_big_rows = std::max(fwd.max_big_rows, bwd.max_big_rows);
        o.max_levels = std::max(fwd.max_levels, bwd.max_levels);
        o.shift_trials = shift_trials;
        o.actual_shift = actual_shift;
        o.setup_seconds = setup_seconds;
        o.device_bytes = A_off.bytes + A_val.bytes + A_col.bytes + A_row.bytes + perm.bytes + rev.bytes +
                         invd.bytes + fwd.bytes() + bwd.bytes();
        o.factor_bytes = fwd.stream_bytes;
        o.spmv_bytes = spmv_bytes;
        o.walk_overlap = ctx ? overlap : 0;
    }

    void one_system(const char* who) const
    {
        if (count != 1)
            throw SsaError(SSA_ERR_ARGUMENT, std::string(who) + ": not available for a batched handle");
    }
    void ordering(int32_t* out) const override
    {
        one_system("ssa_ic_ordering");
        std::copy(h_ordering.begin(), h_ordering.end(), out);
    }
    void schedule(int32_t* co, int32_t* bo) const override
    {
        one_system("ssa_ic_schedule");
        if (co) std::copy(h_color_offsets.begin(), h_color_offsets.end(), co);
        if (bo) std::copy(h_block_offsets.begin(), h_block_offsets.end(), bo);
    }
    void factor(int32_t* rp, int32_t* ci, void* vals, void* iv) const override
    {
        one_system("ssa_ic_factor");
        if (rp) std::copy(h_Lrp.begin(), h_Lrp.end(), rp);
        if (ci) std::copy(h_Lci.begin(), h_Lci.end(), ci);
        if (vals) std::memcpy(vals, h_Lv.data(), h_Lv.size() * sizeof(HT));
        if (iv) std::memcpy(iv, h_invd.data(), h_invd.size() * sizeof(HT));
    }
};


// the factory of one precision pair: T vectors, F factor, HT host scalar
template <typename T, typename F, typename HT>
IcBase* make_impl(ssa_context* ctx, int count, const int32_t* n, const int32_t* const* rp,
                  const int32_t* const* ci, const void* const* values,
                  const sparsesolv_amd::IcSettings& st)
{
    std::vector<const HT*> v(static_cast<size_t>(count));
    for (int q = 0; q < count; ++q) v[size_t(q)] = static_cast<const HT*>(values[q]);
    return new IcImpl<T, F>(ctx, count, n, rp, ci, v.data(), st);
}

}  // namespace ssa_impl
