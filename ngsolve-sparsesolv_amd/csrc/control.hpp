// control.hpp -- scalar control of the Krylov iteration on the device.
//
// The scalar steps of the reference iteration (alpha with its breakdown test,
// the convergence / divergence check, beta and the loop index) run in one
// thread after a deterministic sum of per-workgroup partials (index order,
// fixed 256-thread reduction tree).  Inside the iteration they are fused
// into the kernel that consumes their result; the set-up steps (normaliser,
// first rz) use one-workgroup control kernels with the same reduction.
#pragma once

#include <cfloat>

#include "device_math.hpp"
#include "solver_state.hpp"

namespace ssa {

constexpr int kCtrlThreads = 256;

__device__ inline double dabs(double a) { return fabs(a); }
__device__ inline double dabs(c128 a) { return hypot(a.x, a.y); }
__device__ inline double ddiv(double a, double b) { return a / b; }
__device__ inline c128 ddiv(c128 a, c128 b)
{
    // Smith's algorithm
    if (fabs(b.y) <= fabs(b.x)) {
        const double r = b.y / b.x, d = b.x + b.y * r;
        return c128{(a.x + a.y * r) / d, (a.y - a.x * r) / d};
    }
    const double r = b.x / b.y, d = b.y + b.x * r;
    return c128{(a.x * r + a.y) / d, (a.y * r - a.x) / d};
}

// IterativeSolver::check_convergence (solvers/iterative_solver.hpp:217-249);
// with prm.ngs the COCRSolverNGS test res/rhs_norm <= tol
// (ngsolve/sparsesolv_solvers.hpp:338-347).
// `recheck`: the check of a breakdown, on the residual the last check saw;
// it leaves `improved` alone unless it improves (only before any check),
// so a kernel that reads `improved` while its workgroup 0 runs this step
// reads an unchanging value.
__device__ inline bool check_convergence(SolverState* s, const SolveParams& prm, double nr,
                                         double* hist, bool recheck = false)
{
    const double rel = nr / s->normalizer;
    if (prm.save_history) hist[s->hist_len++] = rel;
    if (!recheck) s->improved = 0;
    if (prm.save_best && rel < s->best_residual) {
        s->best_residual = rel;
        s->improved = 1;
    }
    if (prm.ngs) return rel <= s->tol;
    if (rel < s->tol || nr < prm.abs_tolerance) return true;
    if (prm.divergence_check) {
        if (rel >= s->best_residual * prm.divergence_threshold) ++s->bad_count;
        else s->bad_count = 0;
    }
    return false;
}

__device__ inline void stop(SolverState* s, int status, int converged, int iters, double nr)
{
    s->status = status;
    s->converged = converged;
    s->result_iters = iters;
    s->final_norm = nr;
}

// CG: alpha = rz / pAp (cg_solver.hpp:100-110); COCR: alpha = rho / mu
// (cocr_solver.hpp:104-116); breakdown below 1e-30.
template <typename T>
__device__ inline void ctrl_alpha_op(SolverState* st, const SolveParams& prm, double* hist, T d)
{
    to_pair(d, st->denom);
    if (dabs(d) < 1e-30) {
        const double nr = st->norm_r;
        if (prm.ngs) {
            stop(st, kBreakdown, 0, st->iter, nr);
            return;
        }
        const bool conv = check_convergence(st, prm, nr, hist, /*recheck=*/true);
        stop(st, kBreakdown, conv ? 1 : 0, conv ? st->iter + 1 : st->iter, nr);
        return;
    }
    const T a = ddiv(from_pair<T>(st->rz), d);
    to_pair(a, st->alpha);
}

// after the residual update: ||r||, check_convergence, is_diverging
__device__ inline void ctrl_check_op(SolverState* st, const SolveParams& prm, double* hist,
                                     double s)
{
    const double nr = sqrt(s);
    st->norm_r = nr;
    if (check_convergence(st, prm, nr, hist)) {
        stop(st, kConverged, 1, st->iter + 1, nr);
        return;
    }
    if (!prm.ngs && prm.divergence_check && st->bad_count >= prm.divergence_count)
        stop(st, kDiverged, 0, st->iter + 1, nr);
}

// CG: beta = rz_new / rz_old (cg_solver.hpp:133-135);
// COCR: rho breakdown test on the old rho, beta = rho_new / rho
// (cocr_solver.hpp:149-155).  Then the loop index advances.
template <typename T>
__device__ inline void ctrl_beta_op(SolverState* st, const SolveParams& prm, T v, T old)
{
    if (prm.cocr && dabs(old) < 1e-30) {
        stop(st, kRhoBreakdown, 0, prm.ngs ? st->iter : st->iter + 1, st->norm_r);
        return;
    }
    to_pair(ddiv(v, old), st->beta);
    to_pair(v, st->rz);
    st->iter += 1;
    if (st->iter >= st->maxit) stop(st, kMaxIter, 0, st->maxit, st->norm_r);
}

// Thread t adds partials t, t + B, t + 2B, ... (B threads) in index order,
// then block_sum combines the threads.  Eight of a thread's partials are
// loaded before they are added, so their loads are in flight together (the
// additions, and so the bits, are the same as one at a time).
template <typename T>
__device__ inline T sum_partials(const T* __restrict__ p, int np, T* red)
{
    constexpr int U = 8;
    const int B = int(blockDim.x);
    T s = zero<T>();
    int i = int(threadIdx.x);
    for (; i + (U - 1) * B < np; i += U * B) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = p[i + u * B];
#pragma unroll
        for (int u = 0; u < U; ++u) s = add(s, v[u]);
    }
    for (; i < np; i += B) s = add(s, p[i]);
    return block_sum(s, red);
}

// ---------------------------------------------------------------------------
// Control steps fused into the consuming kernel
//
// A kernel that needs alpha or beta sums the producing kernel's partials
// itself, in every workgroup, in the same order, so all workgroups hold the
// same bits.  Workgroup 0 alone writes the solver state.  No workgroup reads
// a state field that is written in the same kernel: alpha reads rz,
// the check copies rz to rz_prev, beta reads rz_prev and writes rz.
// ---------------------------------------------------------------------------
struct Ctrl {
    SolverState* st;
    double* hist;
    const void* part;      // the producing kernel's partials
    int np;                // their count (per right-hand side)
    int src_spmv;          // batches: the producer ran on the SpMV grid
    SolveParams prm;
    int part_off;          // first partial of the workgroup's state
    int first;             // the workgroup is its state's first
};

// Several independent systems in one launch (null pointers: one system).
// System s owns the rows [row0[s], row0[s] + nrows[s]); on the vector grid
// its workgroups are [vec_g0[s], vec_g0[s+1]) and cover its rows exactly
// as a one-system launch of that many workgroups would (same partials,
// same bits); on the SpMV grid its workgroups are [spmv_g0[s], ...).
struct Systems {
    const int* vec_sys;
    const int* vec_g0;
    const int* spmv_sys;
    const int* spmv_g0;
    const int* row0;
    const int* nrows;
    int hist_stride;
};

// The rows a vector-grid workgroup owns: i = local*kBlock + lane, stepping
// by G*kBlock while i < n, at device row row0 + i; `sys` is its state and
// `slot` the index of its partial.
struct Span {
    index_t row0, n, local, G;
    int sys, slot;
};

__device__ inline Span vec_span(const Systems& y, index_t n)
{
    if (!y.vec_sys) return Span{0, n, index_t(blockIdx.x), index_t(gridDim.x), 0, int(blockIdx.x)};
    const int s = y.vec_sys[blockIdx.x];
    const int g0 = y.vec_g0[s];
    return Span{y.row0[s], y.nrows[s], index_t(blockIdx.x) - g0, index_t(y.vec_g0[s + 1] - g0), s,
                int(blockIdx.x)};
}

// the control step as seen from a workgroup of state `sys` whose index
// among that state's workgroups is `local`
__device__ inline Ctrl ctrl_for(Ctrl c, const Systems& y, int sys, index_t local)
{
    c.first = local == 0;
    c.st += sys;
    if (c.hist) c.hist += size_t(sys) * size_t(y.hist_stride);
    if (!y.vec_sys) {
        c.part_off = sys * c.np;
        return c;
    }
    const int* g0 = c.src_spmv ? y.spmv_g0 : y.vec_g0;
    c.part_off = g0[sys];
    c.np = g0[sys + 1] - g0[sys];
    return c;
}

// sum of np partials, valid in every thread
template <typename T>
__device__ inline T sum_partials_all(const T* __restrict__ p, int np, T* red)
{
    T v = sum_partials(p, np, red);
    if (threadIdx.x == 0) red[0] = v;
    __syncthreads();
    v = red[0];
    __syncthreads();
    return v;
}

// alpha from the pAp (CG) or mu (COCR) partials; false on breakdown (the
// solve is stopped by workgroup 0)
template <typename T>
__device__ inline bool ctrl_alpha_all(const Ctrl& c, T* red, T& alpha)
{
    const T d = sum_partials_all(static_cast<const T*>(c.part) + c.part_off, c.np, red);
    if (c.first && threadIdx.x == 0) ctrl_alpha_op(c.st, c.prm, c.hist, d);
    if (dabs(d) < 1e-30) return false;
    alpha = ddiv(from_pair<T>(c.st->rz), d);
    return true;
}

// convergence check from the |r|^2 partials, in the system's first
// workgroup only; every thread of the kernel must call it (it synchronises
// that workgroup)
__device__ inline void ctrl_check_wg0(const Ctrl& c, double* red)
{
    if (!c.first) return;
    const double s = sum_partials(static_cast<const double*>(c.part) + c.part_off, c.np, red);
    if (threadIdx.x != 0) return;
    c.st->rz_prev[0] = c.st->rz[0];
    c.st->rz_prev[1] = c.st->rz[1];
    ctrl_check_op(c.st, c.prm, c.hist, s);
}

// beta from the new rz / rho partials; false on the COCR rho breakdown
template <typename T>
__device__ inline bool ctrl_beta_all(const Ctrl& c, T* red, T& beta)
{
    const T v = sum_partials_all(static_cast<const T*>(c.part) + c.part_off, c.np, red);
    const T old = from_pair<T>(c.st->rz_prev);
    if (c.first && threadIdx.x == 0) ctrl_beta_op(c.st, c.prm, v, old);
    if (c.prm.cocr && dabs(old) < 1e-30) return false;
    beta = ddiv(v, old);
    return true;
}

// ---------------------------------------------------------------------------
// Control kernels: one workgroup of kCtrlThreads
// ---------------------------------------------------------------------------

// prepare_iteration (iterative_solver.hpp:135-180): normaliser, initial
// history entry after r = b - Ax.  One workgroup per state; pp (optional)
// holds each state's tolerance, iteration budget and whether it runs.
// (static: every translation unit that launches it has its own copy)
static __global__ __launch_bounds__(kCtrlThreads) void k_ctrl_init(const double* __restrict__ pb,
                                                            const double* __restrict__ pr,
                                                            int np, SolverState* st,
                                                            SolveParams prm,
                                                            double* __restrict__ hist,
                                                            Systems y,
                                                            const PassParams* __restrict__ pp)
{
    // the state's partials are those of its vector grid
    const int s = blockIdx.x;
    if (y.vec_sys) {
        const int g0 = y.vec_g0[s];
        pb += g0;
        pr += g0;
        np = y.vec_g0[s + 1] - g0;
    } else {
        pb += size_t(s) * size_t(np);
        pr += size_t(s) * size_t(np);
    }
    st += s;
    if (hist) hist += size_t(s) * size_t(y.hist_stride);
    __shared__ double red[kCtrlThreads / 64];
    const double sb = sum_partials(pb, np, red);
    const double sr = sum_partials(pr, np, red);
    if (threadIdx.x != 0) return;
    const double norm_b = sqrt(sb), init = sqrt(sr);
    double nz = prm.norm_type == 0 ? norm_b : prm.norm_type == 1 ? init : prm.custom_norm;
    if (prm.ngs) {
        // COCRSolverNGS: rhs_norm of the projected rhs, early exit below 1e-30
        nz = norm_b;
    } else if (nz < 1e-15) {
        nz = 1.0;
    }
    st->normalizer = nz;
    st->norm_r = init;
    st->best_residual = DBL_MAX;
    st->bad_count = 0;
    st->improved = 0;
    st->iter = 0;
    st->hist_len = 0;
    st->replace_norm = init;
    st->replace_iter = 0;
    st->status = kRunning;
    st->converged = 0;
    st->tol = pp ? pp[s].tolerance : prm.tolerance;
    st->maxit = pp ? pp[s].max_iterations : prm.max_iterations;
    if (pp && !pp[s].active) {
        stop(st, kConverged, 1, 0, init);
        return;
    }
    if (prm.save_history) hist[st->hist_len++] = prm.ngs ? 1.0 : init / nz;
    if (prm.ngs && norm_b < 1e-30) {
        stop(st, kConverged, 1, 0, 0.0);
        st->normalizer = 1.0;
        return;
    }
    if (st->maxit <= 0) stop(st, kMaxIter, 0, st->maxit, init);
}

// rz_old (CG) or rho (COCR) from dot partials into slot 0 (rz) or 1 (denom);
// one workgroup per state
template <typename T>
__global__ __launch_bounds__(kCtrlThreads) void k_ctrl_store(const T* __restrict__ part, int np,
                                                            SolverState* st, int slot,
                                                            Systems y, int src_spmv)
{
    const int s = blockIdx.x;
    if (y.vec_sys) {
        const int* g0 = src_spmv ? y.spmv_g0 : y.vec_g0;
        part += g0[s];
        np = g0[s + 1] - g0[s];
    } else {
        part += size_t(s) * size_t(np);
    }
    st += s;
    __shared__ T red[kCtrlThreads / 64];
    const T v = sum_partials(part, np, red);
    if (threadIdx.x != 0) return;
    double pr[2];
    to_pair(v, pr);
    double* dst = slot == 0 ? st->rz : st->denom;
    dst[0] = pr[0];
    dst[1] = pr[1];
}

// Mixed-precision COCR, residual replacement: whether state q recomputes
// rt = M^{-1} r now -- when ||r|| has fallen below `drop` times its value at
// the last replacement (or the pass start), or `gap` iterations have passed
// since.  rst[q].status is kRunning for the states that replace and nonzero
// for the others (the replacement's kernels read only the status: the
// others, and stopped states, skip them).
static __global__ __launch_bounds__(64) void k_ctrl_replace(SolverState* __restrict__ st,
                                                            SolverState* __restrict__ rst,
                                                            int count, double drop, int gap)
{
    const int q = int(blockIdx.x) * 64 + int(threadIdx.x);
    if (q >= count) return;
    SolverState* s = st + q;
    int go = 0;
    if (s->status == kRunning) {
        go = s->norm_r <= s->replace_norm * drop || s->iter - s->replace_iter >= gap;
        if (go) {
            s->replace_norm = s->norm_r;
            s->replace_iter = s->iter;
        }
    }
    rst[q].status = go ? kRunning : kMaxIter;
}

}  // namespace ssa
