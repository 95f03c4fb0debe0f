// host_setup.cpp -- host half of ssa_ic_create: ABMC schedule, IC(0)
// factor and HBM layouts (compiled by g++ with -ffp-contract=off).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <exception>
#include <stdexcept>
#include <string>
#include <thread>

#include "host_plan.hpp"

namespace ssa_host {

template <typename F, typename T>
static F narrow(const T& v) { return static_cast<F>(v); }
template <>
cf narrow<cf, cd>(const cd& v) { return cf(float(v.real()), float(v.imag())); }

// Rows moved to ord[row], columns renamed to ord[col], each row's entries
// kept in their original order (so the diagonal stays where the sweep
// expects it and every row sums its terms in the reference's order).
template <typename T>
static sparsesolv_amd::CsrHost<T> relabel(const sparsesolv_amd::CsrHost<T>& M,
                                          const std::vector<int>& ord)
{
    const int n = M.n;
    std::vector<int> inv(n);
    for (int i = 0; i < n; ++i) inv[ord[i]] = i;
    sparsesolv_amd::CsrHost<T> R;
    R.n = n;
    R.rp.assign(size_t(n) + 1, 0);
    for (int r = 0; r < n; ++r) R.rp[r + 1] = R.rp[r] + (M.rp[inv[r] + 1] - M.rp[inv[r]]);
    R.ci.resize(M.ci.size());
    R.v.resize(M.v.size());
    for (int r = 0; r < n; ++r) {
        int pos = R.rp[r];
        for (int k = M.rp[inv[r]]; k < M.rp[inv[r] + 1]; ++k, ++pos) {
            R.ci[pos] = ord[M.ci[k]];
            R.v[pos] = M.v[k];
        }
    }
    return R;
}

// Level schedule of the natural-order factor (core/level_schedule.hpp:58-89):
// level(i) = 1 + max level of the columns j < i of row i of L.  Rows are
// renumbered level by level (natural order inside a level); every level is a
// "colour" of one-row blocks, so the colour sweeps run it unchanged.  The
// backward sweep takes the same levels in reverse: every row that row i of
// U = L^T depends on has a higher forward level.
static void level_schedule(const sparsesolv_amd::CsrHost<double>* Lr,
                           const sparsesolv_amd::CsrHost<cd>* Lc, int n,
                           sparsesolv_amd::AbmcSchedule& S)
{
    const std::vector<int>& rp = Lr ? Lr->rp : Lc->rp;
    const std::vector<int>& ci = Lr ? Lr->ci : Lc->ci;
    std::vector<int> level(n, 0);
    int nlev = 0;
    for (int i = 0; i < n; ++i) {
        int m = -1;
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] < i) m = std::max(m, level[ci[k]]);
        level[i] = m + 1;
        nlev = std::max(nlev, level[i] + 1);
    }
    std::vector<int> count(size_t(nlev) + 1, 0);
    for (int i = 0; i < n; ++i) ++count[level[i] + 1];
    for (int l = 0; l < nlev; ++l) count[l + 1] += count[l];
    S.color_offsets = count;                  // blocks are single rows
    S.block_offsets.resize(size_t(n) + 1);
    for (int i = 0; i <= n; ++i) S.block_offsets[i] = i;
    S.ordering.assign(n, 0);
    S.reverse_ordering.assign(n, 0);
    std::vector<int> next(count.begin(), count.end() - 1);
    for (int i = 0; i < n; ++i) {
        const int r = next[level[i]]++;
        S.ordering[i] = r;
        S.reverse_ordering[r] = i;
    }
}

template <typename T>
static const sparsesolv_amd::CsrHost<double>* as_real(const sparsesolv_amd::CsrHost<T>&) { return nullptr; }
template <>
const sparsesolv_amd::CsrHost<double>* as_real<double>(const sparsesolv_amd::CsrHost<double>& M) { return &M; }
template <typename T>
static const sparsesolv_amd::CsrHost<cd>* as_complex(const sparsesolv_amd::CsrHost<T>&) { return nullptr; }
template <>
const sparsesolv_amd::CsrHost<cd>* as_complex<cd>(const sparsesolv_amd::CsrHost<cd>& M) { return &M; }

template <typename T, typename F>
void build_plan(HostPlan<T, F>& plan, int n, const int* row_ptr, const int* col_idx,
                const T* values, const sparsesolv_amd::IcSettings& settings)
{
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    plan.fac.build(n, row_ptr, col_idx, values, settings, /*keep_permuted_matrix=*/true);
    // the device works in the schedule's row order: ABMC as built, or the
    // level order of the natural-order factor
    const sparsesolv_amd::CsrHost<T>* Ldev = &plan.fac.L;
    const sparsesolv_amd::CsrHost<T>* Udev = &plan.fac.U;
    sparsesolv_amd::CsrHost<T> Lp, Up;
    std::vector<T> invd_dev = plan.fac.inv_diag;
    int block_size = settings.abmc_block_size;
    if (!settings.use_abmc) {
        level_schedule(as_real(plan.fac.L), as_complex(plan.fac.L), n, plan.fac.sched);
        const std::vector<int>& ord = plan.fac.sched.ordering;
        Lp = relabel(plan.fac.L, ord);
        Up = relabel(plan.fac.U, ord);
        plan.fac.A_perm = relabel(plan.fac.A_perm, ord);
        for (int i = 0; i < n; ++i) invd_dev[ord[i]] = plan.fac.inv_diag[i];
        Ldev = &Lp;
        Udev = &Up;
        block_size = 1;
        if (settings.factor_kind == sparsesolv_amd::kFactorSgsMrtr) {
            plan.lcsr = Lp;
            plan.split_d.resize(size_t(n));
            plan.split_id.resize(size_t(n));
            for (int i = 0; i < n; ++i) {
                plan.split_d[size_t(ord[i])] = plan.fac.split_d[size_t(i)];
                plan.split_id[size_t(ord[i])] = plan.fac.split_id[size_t(i)];
            }
        }
    }
    // SpMV slices in the original row order (ordering[i]: device row of
    // original row i)
    // the sweep layouts are built from the factor converted to F
    sparsesolv_amd::CsrHost<F> L, U;
    L.n = U.n = n;
    L.rp = Ldev->rp; L.ci = Ldev->ci;
    U.rp = Udev->rp; U.ci = Udev->ci;
    L.v.resize(Ldev->v.size());
    U.v.resize(Udev->v.size());
    for (size_t k = 0; k < L.v.size(); ++k) L.v[k] = narrow<F>(Ldev->v[k]);
    for (size_t k = 0; k < U.v.size(); ++k) U.v[k] = narrow<F>(Udev->v[k]);
    plan.invd.resize(n);
    for (int i = 0; i < n; ++i) plan.invd[i] = narrow<F>(invd_dev[i]);
    // diagonal_scaling: the factor is of S A S, S = diag(1/sqrt|a_ii|) in
    // the factor's order (ic_preconditioner.hpp:1120-1149); device order as
    // D^-1 above
    if (settings.diagonal_scaling) {
        plan.scale.assign(size_t(n), 1.0);
        const std::vector<int>& ord = plan.fac.sched.ordering;
        for (int i = 0; i < n; ++i)
            plan.scale[settings.use_abmc ? size_t(i) : size_t(ord[i])] = plan.fac.scaling[size_t(i)];
    }
    // the three layouts are independent: built on three threads (one when
    // the caller's thread cap says so)
    if (sparsesolv_amd::setup_threads() < 3) {
        plan.spmv.build(plan.fac.A_perm, plan.fac.sched.ordering);
        plan.fwd.build(plan.fac.sched, L, /*forward=*/true, block_size, plan.invd);
        plan.bwd.build(plan.fac.sched, U, /*forward=*/false, block_size, plan.invd);
        plan.seconds = std::chrono::duration<double>(clk::now() - t0).count();
        return;
    }
    std::exception_ptr err[2];
    std::thread tf([&] {
        try {
            plan.fwd.build(plan.fac.sched, L, /*forward=*/true, block_size, plan.invd);
        } catch (...) {
            err[0] = std::current_exception();
        }
    });
    std::thread tb([&] {
        try {
            plan.bwd.build(plan.fac.sched, U, /*forward=*/false, block_size, plan.invd);
        } catch (...) {
            err[1] = std::current_exception();
        }
    });
    std::exception_ptr err_spmv;
    try {
        plan.spmv.build(plan.fac.A_perm, plan.fac.sched.ordering);
    } catch (...) {
        err_spmv = std::current_exception();
    }
    tf.join();
    tb.join();
    for (const auto& e : {err_spmv, err[0], err[1]})
        if (e) std::rethrow_exception(e);
    plan.seconds = std::chrono::duration<double>(clk::now() - t0).count();
}

template <typename T, typename F>
void build_batch(BatchPlan<T, F>& bp, int count, const int* n, const int* const* row_ptr,
                 const int* const* col_idx, const T* const* values,
                 const sparsesolv_amd::IcSettings& settings)
{
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    if (count < 1) throw std::invalid_argument("batch: count must be >= 1");
    if (count > 1 && settings.factor_kind == sparsesolv_amd::kFactorSgsMrtr)
        throw std::invalid_argument("batch: SGS-MRTR runs one system per handle");
    bp = BatchPlan<T, F>{};
    bp.count = count;
    std::vector<HostPlan<T, F>> plans(static_cast<size_t>(count));
    // independent systems: planned on host threads, errors re-raised in order
    std::vector<std::string> errors(static_cast<size_t>(count));
    // largest systems first (by entries), so the last to finish is a small
    // one; the plans keep the caller's order
    std::vector<int> queue(static_cast<size_t>(count));
    for (int s = 0; s < count; ++s) queue[size_t(s)] = s;
    std::stable_sort(queue.begin(), queue.end(), [&](int a, int b) {
        return row_ptr[a][n[a]] > row_ptr[b][n[b]];
    });
    std::atomic<int> next{0};
    auto work = [&] {
        // the systems run side by side: each plans on one thread (the cap
        // is restored, the calling thread runs this too)
        int& cap = sparsesolv_amd::setup_thread_cap();
        const int saved = cap;
        if (count > 1) cap = 1;
        for (int k; (k = next.fetch_add(1)) < count;) {
            const int s = queue[size_t(k)];
            try {
                sparsesolv_amd::IcSettings own = settings;
                if (settings.batch_block_size) own.abmc_block_size = settings.batch_block_size[s];
                if (settings.batch_num_colors) own.abmc_num_colors = settings.batch_num_colors[s];
                own.batch_block_size = own.batch_num_colors = nullptr;
                build_plan<T, F>(plans[s], n[s], row_ptr[s], col_idx[s], values[s], own);
            } catch (const std::exception& e) {
                errors[s] = e.what();
            }
        }
        cap = saved;
    };
    const int nthreads = std::max(1, std::min<int>(count, sparsesolv_amd::setup_threads()));
    std::vector<std::thread> pool;
    for (int k = 1; k < nthreads; ++k) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    for (int s = 0; s < count; ++s)
        if (!errors[s].empty()) throw std::runtime_error(errors[s]);

    for (const auto& P : plans) {
        bp.actual_shift.push_back(P.fac.actual_shift);
        bp.shift_trials.push_back(P.fac.shift_trials);
        bp.seconds.push_back(P.seconds);
        bp.num_colors.push_back(P.fac.sched.num_colors());
        bp.num_blocks.push_back(P.fac.sched.num_blocks());
        bp.nnz.push_back(int(P.fac.A_perm.ci.size()));
        bp.nnz_l.push_back(int(P.fac.L.v.size()));
    }
    bp.in_off.assign(1, 0);
    for (int s = 0; s < count; ++s) {
        bp.nrows.push_back(n[s]);
        bp.in_off.push_back(bp.in_off.back() + n[s]);
    }
    if (count == 1) {
        HostPlan<T, F>& P = plans[0];
        bp.row0 = {0, n[0]};
        bp.gather_in = P.fac.sched.reverse_ordering;
        bp.gather_out = P.fac.sched.ordering;
        // the layouts move; the plan keeps its factor for the queries
        bp.spmv = std::move(P.spmv);
        bp.fwd = std::move(P.fwd);
        bp.bwd = std::move(P.bwd);
        bp.invd = std::move(P.invd);
        bp.scale = std::move(P.scale);
        bp.lcsr = std::move(P.lcsr);
        bp.split_d = std::move(P.split_d);
        bp.split_id = std::move(P.split_id);
        bp.single = std::move(P);
    } else {
        std::vector<const sparsesolv_amd::SpmvLayout<T>*> sp;
        std::vector<const sparsesolv_amd::SweepLayout<F>*> lf, lb;
        for (const auto& P : plans) {
            sp.push_back(&P.spmv);
            lf.push_back(&P.fwd);
            lb.push_back(&P.bwd);
        }
        // the three layouts side by side (each also copies its systems on
        // threads of its own: much of a combine is first-touch of new memory)
        bp.row0 = sparsesolv_amd::combined_rows(sp, 4);
        std::exception_ptr err_f, err_b;
        std::thread tf([&] {
            try {
                bp.fwd = sparsesolv_amd::SweepLayout<F>::combine(lf, bp.row0, bp.slice_sys_f, bp.walk_sys_f);
            } catch (...) {
                err_f = std::current_exception();
            }
        });
        std::thread tb([&] {
            try {
                bp.bwd = sparsesolv_amd::SweepLayout<F>::combine(lb, bp.row0, bp.slice_sys_b, bp.walk_sys_b);
            } catch (...) {
                err_b = std::current_exception();
            }
        });
        std::vector<int> rows;
        bp.spmv = sparsesolv_amd::combine_spmv(sp, 4, rows);
        tf.join();
        tb.join();
        if (err_f) std::rethrow_exception(err_f);
        if (err_b) std::rethrow_exception(err_b);
        const int N = bp.row0.back();
        bp.invd.assign(size_t(N), F(0));
        bp.gather_in.assign(size_t(N), -1);
        bp.gather_out.assign(size_t(bp.in_off.back()), 0);
        for (int s = 0; s < count; ++s) {
            const auto& S = plans[s].fac.sched;
            for (int i = 0; i < n[s]; ++i) {
                bp.invd[bp.row0[s] + i] = plans[s].invd[i];
                bp.gather_in[bp.row0[s] + i] = bp.in_off[s] + S.reverse_ordering[i];
                bp.gather_out[bp.in_off[s] + i] = bp.row0[s] + S.ordering[i];
            }
        }
        if (settings.diagonal_scaling) {
            // each system's S in its rows; padding rows stay zero
            bp.scale.assign(size_t(N), 0.0);
            for (int s = 0; s < count; ++s)
                std::copy(plans[s].scale.begin(), plans[s].scale.end(), bp.scale.begin() + bp.row0[s]);
        }
    }
    bp.wall_seconds = std::chrono::duration<double>(clk::now() - t0).count();
}

template void build_plan<double, double>(HostPlan<double, double>&, int, const int*,
                                         const int*, const double*,
                                         const sparsesolv_amd::IcSettings&);
template void build_plan<double, float>(HostPlan<double, float>&, int, const int*,
                                        const int*, const double*,
                                        const sparsesolv_amd::IcSettings&);
template void build_plan<cd, cd>(HostPlan<cd, cd>&, int, const int*, const int*, const cd*,
                                 const sparsesolv_amd::IcSettings&);
template void build_plan<cd, cf>(HostPlan<cd, cf>&, int, const int*, const int*, const cd*,
                                 const sparsesolv_amd::IcSettings&);

#define SSA_DEFINE_BATCH(T, F)                                                                 \
    template void build_batch<T, F>(BatchPlan<T, F>&, int, const int*, const int* const*,      \
                                    const int* const*, const T* const*,                        \
                                    const sparsesolv_amd::IcSettings&);
SSA_DEFINE_BATCH(double, double)
SSA_DEFINE_BATCH(double, float)
SSA_DEFINE_BATCH(cd, cd)
SSA_DEFINE_BATCH(cd, cf)
#undef SSA_DEFINE_BATCH

}  // namespace ssa_host
