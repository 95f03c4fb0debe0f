// ic_factor.hpp -- host side of the IC(0) preconditioner setup (C++17).
//
// Follows ICPreconditioner::setup on its ABMC branch
// (preconditioners/ic_preconditioner.hpp:69-201):
//   reorder A by the ABMC permutation, rows sorted by new column (:925-989);
//   keep the lower triangle with the diagonal last (:1075-1115);
//   factor in place, L_ij -= sum_{k<j} L_ik L_jk / d_k and
//   d_i = shift*a_ii - sum_{k<i} L_ik^2 / d_k, unconjugated for complex
//   (:1287-1417), with the auto-shift restart (shift += inc, inc *= 2 after
//   a colour in which some |d_i| < min_diagonal_threshold with a_ii != 0)
//   and the pivot clamp |d_i| <= zero_diagonal_replacement -> +replacement;
//   transpose L into L^T (core/sparse_matrix_csr.hpp:47-72).
// The level-schedule branch (use_abmc = false) factors the matrix in its
// own ordering and restarts at the failing row (:1161-1273).
//
// The arithmetic is the reference's, term for term and in the same order,
// so a real factor is bit-identical when both are compiled without fused
// multiply-add (this header is compiled with -ffp-contract=off).  The search
// for L(j,k) walks row j against row i (a merge of the two sorted column
// lists on the ABMC branch, a scatter of row i on the level branch) instead
// of the reference's linear search; it visits the same k in the same order.
//
// Host threads: the rows of the permuted matrix, and on the ABMC branch the
// blocks of one colour (independent, as in the reference's colour-parallel
// loop), are split over up to 16 threads.  Every row computes the same
// values in the same order whichever thread runs it.
#pragma once

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "abmc_ordering.hpp"
#include "rcm_ordering.hpp"

namespace sparsesolv_amd {

// POD mirror of sparsesolv::SolverConfig (core/solver_config.hpp:33-159);
// the C-ABI struct ssa_config has the same layout.
struct IcSettings {
    double shift = 1.05;
    bool diagonal_scaling = false;
    bool auto_shift = false;
    double shift_increment = 0.05;
    double max_shift_value = 5.0;
    double min_diagonal_threshold = 1e-6;
    double zero_diagonal_replacement = 1e-10;
    int max_shift_trials = 100;
    bool use_abmc = true;
    int abmc_block_size = 4;
    int abmc_num_colors = 4;
    bool abmc_use_rcm = false;
    // device schedule, no counterpart in the reference: -1 auto, 0 off, 1 on
    // (the straggler walks on a second stream)
    int walk_overlap = -1;
    // device schedule, no counterpart in the reference: right-hand sides per
    // launch of the several-right-hand-side calls (0: 16)
    int rhs_block = 0;
    // batches: per-system ABMC block size and colour count (null: the
    // fields above for every system)
    const int* batch_block_size = nullptr;
    const int* batch_num_colors = nullptr;
    // what the factor is: kFactorIc the IC(0) factorisation; kFactorSgs
    // A's lower triangle with the pivots diag(A), whose sweeps apply the SGS
    // preconditioner (D + L) D^-1 (D + U) (preconditioners/
    // sgs_preconditioner.hpp:22-40); kFactorSgsMrtr the lower triangle of
    // D A D, D = diag(1/sqrt|a_ii|), that SGS-MRTR solves with
    // (solvers/sgs_mrtr_solver.hpp:285-303).  Both split factors are in the
    // natural order (the level path).
    int factor_kind = 0;
};

constexpr int kFactorIc = 0;
constexpr int kFactorSgs = 1;
constexpr int kFactorSgsMrtr = 2;

template <typename T>
inline double abs_of(const T& v) { return std::abs(v); }

template <typename T>
struct CsrHost {
    index_t n = 0;
    std::vector<index_t> rp, ci;
    std::vector<T> v;
};

// Threads of the host setup's parallel loops: up to 16, fewer when the
// calling thread lowers its cap (a pool that already plans several systems
// side by side sets 1 in its workers).
inline int& setup_thread_cap()
{
    static thread_local int cap = 16;
    return cap;
}

inline int setup_threads()
{
    const unsigned h = std::thread::hardware_concurrency();
    return std::max(1, std::min(setup_thread_cap(), int(std::min(16u, std::max(1u, h)))));
}

// fn(lo, hi) over [0, count) in contiguous chunks of at least `grain`, one
// chunk per thread.  An exception of the lowest chunk that throws is
// rethrown (the serial loop's first failure).
template <typename Fn>
inline void parallel_chunks(index_t count, index_t grain, Fn fn)
{
    const index_t nt = std::max<index_t>(
        1, std::min<index_t>(setup_threads(), count / std::max<index_t>(grain, 1)));
    if (nt <= 1) {
        fn(index_t(0), count);
        return;
    }
    std::vector<std::exception_ptr> err(static_cast<size_t>(nt));
    std::vector<std::thread> pool;
    pool.reserve(size_t(nt));
    for (index_t t = 0; t < nt; ++t)
        pool.emplace_back([&, t] {
            try {
                fn(index_t(int64_t(count) * t / nt), index_t(int64_t(count) * (t + 1) / nt));
            } catch (...) {
                err[size_t(t)] = std::current_exception();
            }
        });
    for (auto& th : pool) th.join();
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
}

// fn(i) for every i in [0, count), the items handed out one at a time in
// increasing order to up to setup_threads() threads (items of uneven cost:
// put the costly ones first).  fn must not throw.
template <typename Fn>
inline void parallel_items(index_t count, Fn fn)
{
    const index_t nt = std::max<index_t>(1, std::min<index_t>(setup_threads(), count));
    std::atomic<index_t> next{0};
    auto run = [&] {
        for (index_t i; (i = next.fetch_add(1)) < count;) fn(i);
    };
    std::vector<std::thread> pool;
    pool.reserve(size_t(nt));
    for (index_t t = 1; t < nt; ++t) pool.emplace_back(run);
    run();
    for (auto& th : pool) th.join();
}

// B[perm i][perm j] = A[i][j], each row sorted by column (unique columns).
template <typename T>
CsrHost<T> permute_csr(index_t n, const index_t* rp, const index_t* ci, const T* v,
                       const std::vector<index_t>& perm)
{
    CsrHost<T> out;
    out.n = n;
    out.rp.assign(size_t(n) + 1, 0);
    for (index_t i = 0; i < n; ++i) out.rp[perm[i] + 1] = rp[i + 1] - rp[i];
    for (index_t i = 0; i < n; ++i) out.rp[i + 1] += out.rp[i];
    out.ci.resize(out.rp[n]);
    out.v.resize(out.rp[n]);
    parallel_chunks(n, 4096, [&](index_t lo, index_t hi) {
        std::vector<std::pair<index_t, index_t>> tmp;
        for (index_t i = lo; i < hi; ++i) {
            const index_t ni = perm[i];
            tmp.clear();
            for (index_t k = rp[i]; k < rp[i + 1]; ++k) tmp.emplace_back(perm[ci[k]], k);
            std::sort(tmp.begin(), tmp.end(),
                      [](const auto& a, const auto& b) { return a.first < b.first; });
            index_t pos = out.rp[ni];
            for (const auto& e : tmp) { out.ci[pos] = e.first; out.v[pos] = v[e.second]; ++pos; }
        }
    });
    return out;
}

template <typename T>
struct IcFactorHost {
    index_t n = 0;
    AbmcSchedule sched;                 // empty on the level path
    CsrHost<T> L;                       // lower triangle, diagonal last
    CsrHost<T> U;                       // L^T, diagonal first
    std::vector<T> inv_diag;
    std::vector<double> scaling;        // 1/sqrt|a_ii| if diagonal_scaling
    // kFactorSgsMrtr: D_i = 1/sqrt|a_ii| and sqrt|a_ii| (1 where |a_ii| <=
    // 1e-15), natural order
    std::vector<T> split_d, split_id;
    double actual_shift = 0.0;
    int shift_trials = 0;
    CsrHost<T> A_perm;                  // A in ABMC order (Krylov operator)

    void build(index_t n_, const index_t* rp, const index_t* ci, const T* v,
               const IcSettings& cfg, bool keep_permuted_matrix)
    {
        n = n_;
        if (cfg.factor_kind != kFactorIc) {
            build_split(rp, ci, v, cfg.factor_kind, keep_permuted_matrix);
            return;
        }
        if (cfg.use_abmc) {
            if (cfg.abmc_use_rcm) {
                // ABMC of the RCM-reordered matrix (ic_preconditioner.hpp:97-113):
                // the schedule's ordering becomes the composite original ->
                // RCM -> ABMC one (:588-595), the matrix is permuted by it
                std::vector<index_t> rcm, rcm_rev;
                rcm_ordering(rp, ci, n, rcm, rcm_rev);
                const CsrHost<T> R = permute_csr(n, rp, ci, v, rcm);
                sched.build(R.rp.data(), R.ci.data(), n, cfg.abmc_block_size, cfg.abmc_num_colors);
                std::vector<index_t> comp(static_cast<size_t>(n));
                for (index_t i = 0; i < n; ++i) comp[size_t(i)] = sched.ordering[size_t(rcm[size_t(i)])];
                for (index_t i = 0; i < n; ++i) sched.reverse_ordering[size_t(comp[size_t(i)])] = i;
                sched.ordering = std::move(comp);
            } else {
                sched.build(rp, ci, n, cfg.abmc_block_size, cfg.abmc_num_colors);
            }
            CsrHost<T> P = permute_csr(n, rp, ci, v, sched.ordering);
            extract_lower(P.n, P.rp.data(), P.ci.data(), P.v.data());
            if (keep_permuted_matrix) A_perm = std::move(P);
        } else {
            extract_lower(n, rp, ci, v);
            if (keep_permuted_matrix) {
                A_perm.n = n;
                A_perm.rp.assign(rp, rp + n + 1);
                A_perm.ci.assign(ci, ci + rp[n]);
                A_perm.v.assign(v, v + rp[n]);
            }
        }
        const std::vector<T> original = L.v;
        if (cfg.diagonal_scaling) scaling_factors();
        factor(cfg, original);
        transpose();
    }

private:
    // The SGS and SGS-MRTR factors (no factorisation), natural order.  Every
    // row of L ends with its diagonal entry (a zero one where A stores none)
    // so the sweeps find the strictly lower part in front of it.
    void build_split(const index_t* rp, const index_t* ci, const T* v, int kind,
                     bool keep_permuted_matrix)
    {
        L.n = n;
        L.rp.assign(size_t(n) + 1, 0);
        std::vector<T> diag(static_cast<size_t>(n), T(0));
        std::vector<char> has(static_cast<size_t>(n), 0);
        for (index_t i = 0; i < n; ++i) {
            index_t c = 0;
            for (index_t k = rp[i]; k < rp[i + 1]; ++k) {
                if (ci[k] < i) ++c;
                if (ci[k] == i) { has[i] = 1; diag[i] = v[k]; }
            }
            L.rp[i + 1] = L.rp[i] + c + 1;
        }
        if (kind == kFactorSgsMrtr) {
            // sgs_mrtr_solver.hpp:88-98
            split_d.assign(static_cast<size_t>(n), T(1));
            split_id.assign(static_cast<size_t>(n), T(1));
            for (index_t i = 0; i < n; ++i) {
                if (!has[i])
                    throw std::runtime_error("SGS-MRTR: missing diagonal at row " + std::to_string(i));
                const double a = abs_of(diag[i]);
                if (a > 1e-15) {
                    const double sq = std::sqrt(a);
                    split_d[i] = T(1.0 / sq);
                    split_id[i] = T(sq);
                }
            }
        }
        L.ci.resize(L.rp[n]);
        L.v.resize(L.rp[n]);
        inv_diag.assign(static_cast<size_t>(n), T(1));
        for (index_t i = 0; i < n; ++i) {
            index_t pos = L.rp[i];
            for (index_t k = rp[i]; k < rp[i + 1]; ++k) {
                if (ci[k] >= i) continue;
                L.ci[pos] = ci[k];
                // L of D A D: (D_i a_ij) D_j (:292)
                L.v[pos] = kind == kFactorSgsMrtr ? split_d[i] * v[k] * split_d[ci[k]] : v[k];
                ++pos;
            }
            L.ci[pos] = i;
            if (kind == kFactorSgsMrtr) {
                L.v[pos] = split_d[i] * diag[i] * split_d[i];
                inv_diag[i] = T(1) / L.v[pos];      // the reference divides by L_ii
            } else {
                // sgs_preconditioner.hpp:80-87
                L.v[pos] = diag[i];
                if (abs_of(diag[i]) > 1e-15) inv_diag[i] = T(1) / diag[i];
            }
        }
        if (keep_permuted_matrix) {
            A_perm.n = n;
            A_perm.rp.assign(rp, rp + n + 1);
            A_perm.ci.assign(ci, ci + rp[n]);
            A_perm.v.assign(v, v + rp[n]);
        }
        actual_shift = 0.0;
        shift_trials = 0;
        transpose();
    }

    void extract_lower(index_t nn, const index_t* rp, const index_t* ci, const T* v)
    {
        L.n = nn;
        L.rp.assign(size_t(nn) + 1, 0);
        parallel_chunks(nn, 4096, [&](index_t lo, index_t hi) {
            for (index_t i = lo; i < hi; ++i) {
                index_t c = 0;
                for (index_t k = rp[i]; k < rp[i + 1]; ++k) c += ci[k] <= i;
                L.rp[i + 1] = c;
            }
        });
        for (index_t i = 0; i < nn; ++i) L.rp[i + 1] += L.rp[i];
        L.ci.resize(L.rp[nn]);
        L.v.resize(L.rp[nn]);
        parallel_chunks(nn, 4096, [&](index_t lo, index_t hi) {
            for (index_t i = lo; i < hi; ++i) {
                index_t pos = L.rp[i];
                for (index_t k = rp[i]; k < rp[i + 1]; ++k)
                    if (ci[k] <= i) { L.ci[pos] = ci[k]; L.v[pos] = v[k]; ++pos; }
            }
        });
    }

    void scaling_factors()
    {
        scaling.assign(n, 1.0);
        for (index_t i = 0; i < n; ++i) {
            const index_t d = L.rp[i + 1] - 1;
            if (d >= L.rp[i] && L.ci[d] == i) {
                const double a = abs_of(L.v[d]);
                scaling[i] = a > 0.0 ? 1.0 / std::sqrt(a) : 1.0;
            }
        }
    }

    void scale_into(const std::vector<T>& original)
    {
        for (index_t i = 0; i < n; ++i)
            for (index_t k = L.rp[i]; k < L.rp[i + 1]; ++k)
                L.v[k] = original[k] * static_cast<T>(scaling[i]) *
                         static_cast<T>(scaling[L.ci[k]]);
    }

    // Off-diagonal entries of row i (ABMC branch): L(j,k) is looked up by
    // merging row j with row i, both sorted by column (the permuted matrix
    // sorts every row), which visits the same k in the same order as the
    // scatter of row i in the level branch below.
    void update_row_merge(index_t i)
    {
        const index_t rs = L.rp[i], re = L.rp[i + 1];
        for (index_t kk = rs; kk < re - 1; ++kk) {
            const index_t j = L.ci[kk];
            if (j >= i) break;
            T s = L.v[kk];
            index_t p = rs;
            // terms k < j common to rows i and j, ascending k
            for (index_t jj = L.rp[j]; jj < L.rp[j + 1]; ++jj) {
                const index_t k = L.ci[jj];
                if (k >= j) break;
                while (p < kk && L.ci[p] < k) ++p;
                if (p < kk && L.ci[p] == k) s -= L.v[p] * L.v[jj] * inv_diag[k];
            }
            L.v[kk] = s;
        }
    }

    // Pivot of row i: d_i = shift * a_ii - sum L_ik^2 / d_k, stored with
    // its inverse (clamped).  Returns whether it falls under the auto-shift
    // threshold (the caller decides about the restart).
    bool pivot_row(index_t i, double shift, const IcSettings& cfg, bool clamp_on_breakdown)
    {
        const index_t rs = L.rp[i], re = L.rp[i + 1];
        const index_t dp = re - 1;
        if (dp < rs || L.ci[dp] != i)
            throw std::runtime_error("IC decomposition failed: missing diagonal at row " +
                                     std::to_string(i));
        const T orig = L.v[dp];
        T s = orig * static_cast<T>(shift);
        for (index_t kk = rs; kk < dp; ++kk) {
            const index_t k = L.ci[kk];
            if (k >= i) break;
            s -= L.v[kk] * L.v[kk] * inv_diag[k];
        }
        L.v[dp] = s;
        const double abs_s = abs_of(s);
        const bool low = cfg.auto_shift && abs_s < cfg.min_diagonal_threshold && abs_of(orig) > 0.0;
        if (low && !clamp_on_breakdown) return true;
        if (abs_s > cfg.zero_diagonal_replacement) {
            inv_diag[i] = T(1) / s;
        } else {
            const T safe = static_cast<T>(cfg.zero_diagonal_replacement);
            L.v[dp] = safe;
            inv_diag[i] = T(1) / safe;
        }
        return low;
    }

    void factor(const IcSettings& cfg, const std::vector<T>& original)
    {
        inv_diag.assign(n, T(0));
        double shift = cfg.shift;
        double inc = cfg.shift_increment;
        int trials = 0;
        const bool abmc = sched.built();
        const index_t nc = abmc ? sched.num_colors() : 1;
        bool restart = true;
        while (restart) {
            restart = false;
            if (trials > 0) L.v = original;
            if (cfg.diagonal_scaling) scale_into(original);
            if (abmc) {
                // the blocks of a colour are independent: split over threads;
                // a pivot under the threshold restarts after the colour
                for (index_t c = 0; c < nc && !restart; ++c) {
                    const index_t b0 = sched.color_offsets[c], b1 = sched.color_offsets[c + 1];
                    std::atomic<bool> low{false};
                    parallel_chunks(b1 - b0, 64, [&](index_t lo, index_t hi) {
                        bool any = false;
                        for (index_t b = b0 + lo; b < b0 + hi; ++b)
                            for (index_t i = sched.block_offsets[b]; i < sched.block_offsets[b + 1]; ++i) {
                                update_row_merge(i);
                                any |= pivot_row(i, shift, cfg, true);
                            }
                        if (any) low.store(true);
                    });
                    const bool breakdown = low.load();
                    if (breakdown && shift < cfg.max_shift_value && trials < cfg.max_shift_trials) {
                        shift += inc; inc *= 2; ++trials;
                        restart = true;
                    }
                }
            } else {
                // natural order, one row after the other; the restart comes at
                // the failing row.  where[c] = position of column c in row i,
                // or -1
                std::vector<index_t> where(n, -1);
                for (index_t i = 0; i < n; ++i) {
                    const index_t rs = L.rp[i], re = L.rp[i + 1];
                    for (index_t q = rs; q < re; ++q) where[L.ci[q]] = q;
                    for (index_t kk = rs; kk < re - 1; ++kk) {
                        const index_t j = L.ci[kk];
                        if (j >= i) break;
                        T s = L.v[kk];
                        // terms k < j common to rows i and j, ascending k
                        for (index_t jj = L.rp[j]; jj < L.rp[j + 1]; ++jj) {
                            const index_t k = L.ci[jj];
                            if (k >= j) break;
                            const index_t ii = where[k];
                            if (ii >= 0 && ii < kk)
                                s -= L.v[ii] * L.v[jj] * inv_diag[k];
                        }
                        L.v[kk] = s;
                    }
                    for (index_t q = rs; q < re; ++q) where[L.ci[q]] = -1;
                    if (pivot_row(i, shift, cfg, false)) {
                        if (shift < cfg.max_shift_value && trials < cfg.max_shift_trials) {
                            shift += inc; inc *= 2; ++trials;
                            restart = true;
                            break;
                        }
                        // no restart left: clamp as usual
                        const index_t dp = L.rp[i + 1] - 1;
                        const double abs_s = abs_of(L.v[dp]);
                        if (abs_s > cfg.zero_diagonal_replacement) {
                            inv_diag[i] = T(1) / L.v[dp];
                        } else {
                            const T safe = static_cast<T>(cfg.zero_diagonal_replacement);
                            L.v[dp] = safe;
                            inv_diag[i] = T(1) / safe;
                        }
                    }
                }
            }
        }
        actual_shift = shift;
        shift_trials = trials;
    }

    // U = L^T, each row of U in ascending column order.  A thread owns a
    // range [c0, c1) of U's rows (L's columns) and scans the rows i >= c0 of
    // L (L is lower triangular, rows sorted by column) in ascending order.
    void transpose()
    {
        U.n = n;
        const index_t nnz = L.rp[n];
        U.rp.assign(size_t(n) + 1, 0);
        U.ci.resize(nnz);
        U.v.resize(nnz);
        auto in_range = [&](index_t i, index_t c0) {
            return index_t(std::lower_bound(L.ci.begin() + L.rp[i], L.ci.begin() + L.rp[i + 1], c0) -
                           L.ci.begin());
        };
        parallel_chunks(n, 4096, [&](index_t c0, index_t c1) {
            for (index_t i = c0; i < n; ++i)
                for (index_t k = in_range(i, c0); k < L.rp[i + 1] && L.ci[k] < c1; ++k) U.rp[L.ci[k] + 1]++;
        });
        for (index_t i = 0; i < n; ++i) U.rp[i + 1] += U.rp[i];
        std::vector<index_t> fill(U.rp.begin(), U.rp.end() - 1);
        parallel_chunks(n, 4096, [&](index_t c0, index_t c1) {
            for (index_t i = c0; i < n; ++i)
                for (index_t k = in_range(i, c0); k < L.rp[i + 1] && L.ci[k] < c1; ++k) {
                    const index_t p = fill[L.ci[k]]++;
                    U.ci[p] = i;
                    U.v[p] = L.v[k];
                }
        });
    }
};

}  // namespace sparsesolv_amd
