// block_kernels.hpp -- vector and control kernels for a block of W
// right-hand sides of one system (interleaved rows: element (i, r) at
// v[i*W + r]; the SpMV and the sweeps take W as a template in kernels.hpp).
//
// The kernels that reduce (dots, norms, the residual updates) give each
// thread whole rows: the partial of right-hand side r in workgroup g sums
// rows g*kBlock + t, stepping by G*kBlock, with block_sum's tree -- the same
// rows in the same order as workgroup g of the one-vector kernel, so a block
// solve reproduces the one-vector solves bit for bit.  Element-wise kernels
// take the flat, coalesced mapping (thread e, element e; G*kBlock is a
// multiple of W, so a thread keeps one right-hand side).  The control steps
// run as W one-workgroup kernels between producer and consumer: a fused
// step would have every consuming workgroup re-sum W sets of partials.
// Partials of right-hand side r occupy [r*np, (r+1)*np) for a producer of
// np workgroups.
#pragma once

#include "kernels.hpp"

namespace ssa {

// values of a row moved at once: 64 bytes (four dwordx4) or the whole row
template <typename T, int W>
constexpr int row_chunk() { return int(64 / sizeof(T)) < W ? int(64 / sizeof(T)) : W; }

template <int W>
__device__ inline void running_mask(const SolverState* __restrict__ st, bool (&run)[W])
{
#pragma unroll
    for (int r = 0; r < W; ++r) run[r] = st[r].status == 0;
}

// partial sums of conj?(a_i) b_i, accumulated in T (operands converted)
template <typename T, typename TA, typename TB, int W>
__global__ __launch_bounds__(kBlock) void k_dot_blk(index_t n, const TA* __restrict__ a,
                                                   const TB* __restrict__ b, int conj,
                                                   T* __restrict__ partials)
{
    __shared__ T red[kWavesPerBlock * W];
    T d[W];
#pragma unroll
    for (int r = 0; r < W; ++r) d[r] = zero<T>();
    constexpr int C = row_chunk<TA, W>() < row_chunk<TB, W>() ? row_chunk<TA, W>() : row_chunk<TB, W>();
    for (index_t i = index_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += index_t(gridDim.x) * kBlock) {
#pragma unroll
        for (int c0 = 0; c0 < W; c0 += C) {
            TA av[C];
            TB bv[C];
            load_row<TA, C>(a + size_t(i) * W + c0, av);
            load_row<TB, C>(b + size_t(i) * W + c0, bv);
#pragma unroll
            for (int r = 0; r < C; ++r)
                d[c0 + r] = fma_(conj_if(cvt<T>(av[r]), conj != 0), cvt<T>(bv[r]), d[c0 + r]);
        }
    }
    const T t = block_sum_multi<T, W>(d, red);
    if (threadIdx.x < W) partials[size_t(threadIdx.x) * gridDim.x + blockIdx.x] = t;
}

// partial sums of |a_i|^2 and (optionally) |b_i|^2
template <typename T, int W>
__global__ __launch_bounds__(kBlock) void k_norms_blk(index_t n, const T* __restrict__ a,
                                                     const T* __restrict__ b,
                                                     double* __restrict__ pa,
                                                     double* __restrict__ pb)
{
    __shared__ double red[kWavesPerBlock * W];
    double sa[W], sb[W];
#pragma unroll
    for (int r = 0; r < W; ++r) sa[r] = sb[r] = 0.0;
    constexpr int C = row_chunk<T, W>();
    for (index_t i = index_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += index_t(gridDim.x) * kBlock) {
#pragma unroll
        for (int c0 = 0; c0 < W; c0 += C) {
            T av[C];
            load_row<T, C>(a + size_t(i) * W + c0, av);
#pragma unroll
            for (int r = 0; r < C; ++r) sa[c0 + r] += norm2(av[r]);
            if (b) {
                T bv[C];
                load_row<T, C>(b + size_t(i) * W + c0, bv);
#pragma unroll
                for (int r = 0; r < C; ++r) sb[c0 + r] += norm2(bv[r]);
            }
        }
    }
    double t = block_sum_multi<double, W>(sa, red);
    if (threadIdx.x < W) pa[size_t(threadIdx.x) * gridDim.x + blockIdx.x] = t;
    if (b) {
        t = block_sum_multi<double, W>(sb, red);
        if (threadIdx.x < W) pb[size_t(threadIdx.x) * gridDim.x + blockIdx.x] = t;
    }
}

// CG: x += alpha p; r -= alpha Ap; partial |r|^2; r also in the
// preconditioner's precision when r_low is given (alpha from the state)
template <typename T, typename P, int W>
__global__ __launch_bounds__(kBlock) void k_cg_update_blk(index_t n, T* __restrict__ x,
                                                         T* __restrict__ r, const T* __restrict__ p,
                                                         const T* __restrict__ Ap,
                                                         double* __restrict__ partials,
                                                         const SolverState* __restrict__ st,
                                                         P* __restrict__ r_low)
{
    if (all_stopped<W>(st)) return;
    __shared__ double red[kWavesPerBlock * W];
    bool run[W];
    running_mask<W>(st, run);
    T alpha[W];
    double s[W];
#pragma unroll
    for (int q = 0; q < W; ++q) {
        alpha[q] = from_pair<T>(st[q].alpha);
        s[q] = 0.0;
    }
    constexpr int C = row_chunk<T, W>();
    for (index_t i = index_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += index_t(gridDim.x) * kBlock) {
#pragma unroll
        for (int c0 = 0; c0 < W; c0 += C) {
            const size_t at = size_t(i) * W + c0;
            T xv[C], rv[C], pv[C], av[C];
            P lv[C];
            load_row<T, C>(x + at, xv);
            load_row<T, C>(r + at, rv);
            load_row<T, C>(p + at, pv);
            load_row<T, C>(Ap + at, av);
#pragma unroll
            for (int k = 0; k < C; ++k) {
                const int q = c0 + k;
                if (run[q]) {
                    xv[k] = fma_(alpha[q], pv[k], xv[k]);
                    rv[k] = fms_(alpha[q], av[k], rv[k]);
                    s[q] += norm2(rv[k]);
                }
                lv[k] = cvt<P>(rv[k]);
            }
            store_row<T, C>(x + at, xv);
            store_row<T, C>(r + at, rv);
            if (r_low) store_row<P, C>(r_low + at, lv);
        }
    }
    const double t = block_sum_multi<double, W>(s, red);
    if (threadIdx.x < W) partials[size_t(threadIdx.x) * gridDim.x + blockIdx.x] = t;
}

// COCR: x += alpha p; r -= alpha q; rt -= alpha qt; partial |r|^2 (qt may
// be in the preconditioner's precision)
template <typename T, typename TQ, int W>
__global__ __launch_bounds__(kBlock) void k_cocr_update_blk(index_t n, T* __restrict__ x,
                                                           T* __restrict__ r, T* __restrict__ rt,
                                                           const T* __restrict__ p,
                                                           const T* __restrict__ q,
                                                           const TQ* __restrict__ qt,
                                                           double* __restrict__ partials,
                                                           const SolverState* __restrict__ st)
{
    if (all_stopped<W>(st)) return;
    __shared__ double red[kWavesPerBlock * W];
    bool run[W];
    running_mask<W>(st, run);
    T alpha[W];
    double s[W];
#pragma unroll
    for (int k = 0; k < W; ++k) {
        alpha[k] = from_pair<T>(st[k].alpha);
        s[k] = 0.0;
    }
    constexpr int C = row_chunk<T, W>();
    for (index_t i = index_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += index_t(gridDim.x) * kBlock) {
#pragma unroll
        for (int c0 = 0; c0 < W; c0 += C) {
            const size_t at = size_t(i) * W + c0;
            T xv[C], rv[C], tv[C], pv[C], qv[C];
            TQ qtv[C];
            load_row<T, C>(x + at, xv);
            load_row<T, C>(r + at, rv);
            load_row<T, C>(rt + at, tv);
            load_row<T, C>(p + at, pv);
            load_row<T, C>(q + at, qv);
            load_row<TQ, C>(qt + at, qtv);
#pragma unroll
            for (int k = 0; k < C; ++k) {
                const int j = c0 + k;
                if (run[j]) {
                    xv[k] = fma_(alpha[j], pv[k], xv[k]);
                    rv[k] = fms_(alpha[j], qv[k], rv[k]);
                    tv[k] = fms_(alpha[j], cvt<T>(qtv[k]), tv[k]);
                    s[j] += norm2(rv[k]);
                }
            }
            store_row<T, C>(x + at, xv);
            store_row<T, C>(r + at, rv);
            store_row<T, C>(rt + at, tv);
        }
    }
    const double t = block_sum_multi<double, W>(s, red);
    if (threadIdx.x < W) partials[size_t(threadIdx.x) * gridDim.x + blockIdx.x] = t;
}

// ---- element-wise (flat mapping over the n*W elements) ---------------------

// dst = src converted; right-hand sides that have stopped are left alone
// (st null: all copied)
template <typename TS, typename TD, int W>
__global__ __launch_bounds__(kBlock) void k_copy_blk(size_t total, const TS* __restrict__ src,
                                                    TD* __restrict__ dst,
                                                    const SolverState* __restrict__ st)
{
    const int q = int(threadIdx.x) & (W - 1);
    if (st && st[q].status) return;
    for (size_t e = size_t(blockIdx.x) * kBlock + threadIdx.x; e < total;
         e += size_t(gridDim.x) * kBlock)
        dst[e] = cvt<TD>(src[e]);
}

// CG: best <- x when the last check improved (also after a stop), then
// p = z + beta p (beta from the state)
template <typename T, typename TZ, int W>
__global__ __launch_bounds__(kBlock) void k_xpby_blk(size_t total, const TZ* __restrict__ z,
                                                    T* __restrict__ p, const T* __restrict__ x,
                                                    T* __restrict__ best,
                                                    const SolverState* __restrict__ st)
{
    const int q = int(threadIdx.x) & (W - 1);
    const bool copy = best && st[q].improved;
    const bool run = st[q].status == 0;
    if (!copy && !run) return;
    const T beta = from_pair<T>(st[q].beta);
    for (size_t e = size_t(blockIdx.x) * kBlock + threadIdx.x; e < total;
         e += size_t(gridDim.x) * kBlock) {
        if (copy) best[e] = x[e];
        if (run) p[e] = fma_(beta, p[e], cvt<T>(z[e]));
    }
}

// COCR: best <- x when improved; p = rt + beta p; q = t + beta q (q also in
// the preconditioner's precision when q_low is given)
template <typename T, typename P, int W>
__global__ __launch_bounds__(kBlock) void k_cocr_pq_blk(size_t total, const T* __restrict__ rt,
                                                       const T* __restrict__ t, T* __restrict__ p,
                                                       T* __restrict__ q, const T* __restrict__ x,
                                                       T* __restrict__ best,
                                                       const SolverState* __restrict__ st,
                                                       P* __restrict__ q_low)
{
    const int k = int(threadIdx.x) & (W - 1);
    const bool copy = best && st[k].improved;
    const bool run = st[k].status == 0;
    if (!copy && !run) return;
    const T beta = from_pair<T>(st[k].beta);
    for (size_t e = size_t(blockIdx.x) * kBlock + threadIdx.x; e < total;
         e += size_t(gridDim.x) * kBlock) {
        if (copy) best[e] = x[e];
        if (run) {
            p[e] = fma_(beta, p[e], rt[e]);
            const T qi = fma_(beta, q[e], t[e]);
            q[e] = qi;
            if (q_low) q_low[e] = cvt<P>(qi);
        }
    }
}

// dst (interleaved, n rows) <- the caller's m vectors (vector r at
// src + r*sstride, in its own ordering): dst[i*W + r] = src_r[map[i]];
// columns m..W-1 (block padding) and map -1 are zero
template <typename T, int W>
__global__ __launch_bounds__(kBlock) void k_gather_blk(index_t n, const index_t* __restrict__ map,
                                                      const T* __restrict__ src, size_t sstride,
                                                      int m, T* __restrict__ dst)
{
    const size_t total = size_t(n) * W;
    for (size_t e = size_t(blockIdx.x) * kBlock + threadIdx.x; e < total;
         e += size_t(gridDim.x) * kBlock) {
        const int r = int(e & size_t(W - 1));
        const index_t k = map[e / W];
        dst[e] = (r < m && k >= 0) ? src[size_t(r) * sstride + size_t(k)] : zero<T>();
    }
}

// the caller's m vectors (vector r at dst + r*tot) <- dst_r[i] =
// src[map[i]*W + r]
template <typename T, int W>
__global__ __launch_bounds__(kBlock) void k_scatter_blk(index_t tot, const index_t* __restrict__ map,
                                                       const T* __restrict__ src, int m,
                                                       T* __restrict__ dst)
{
    const size_t total = size_t(tot) * size_t(m);
    for (size_t e = size_t(blockIdx.x) * kBlock + threadIdx.x; e < total;
         e += size_t(gridDim.x) * kBlock) {
        const size_t r = e / size_t(tot), i = e % size_t(tot);
        dst[e] = src[size_t(map[i]) * W + r];
    }
}

// column r of an interleaved vector: dst[i*W + r] = src[i*W + r]
template <typename T>
__global__ __launch_bounds__(kBlock) void k_copy_column(index_t n, int w, int r,
                                                       const T* __restrict__ src,
                                                       T* __restrict__ dst)
{
    for (index_t i = index_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += index_t(gridDim.x) * kBlock)
        dst[size_t(i) * w + r] = src[size_t(i) * w + r];
}

// ---- control steps: one workgroup of kCtrlThreads per right-hand side ------

// alpha from the pAp (CG) or mu (COCR) partials, or the breakdown stop
template <typename T>
__global__ __launch_bounds__(kCtrlThreads) void k_ctrl_alpha_blk(const T* __restrict__ part, int np,
                                                                SolverState* st, SolveParams prm,
                                                                double* hist, int hist_stride)
{
    const int r = blockIdx.x;
    st += r;
    if (st->status) return;
    __shared__ T red[kCtrlThreads / 64];
    const T d = sum_partials(part + size_t(r) * size_t(np), np, red);
    if (threadIdx.x == 0) ctrl_alpha_op(st, prm, hist ? hist + size_t(r) * hist_stride : nullptr, d);
}

// convergence check from the |r|^2 partials
static __global__ __launch_bounds__(kCtrlThreads) void k_ctrl_check_blk(
    const double* __restrict__ part, int np, SolverState* st, SolveParams prm, double* hist,
    int hist_stride)
{
    const int r = blockIdx.x;
    st += r;
    if (st->status) return;
    __shared__ double red[kCtrlThreads / 64];
    const double s = sum_partials(part + size_t(r) * size_t(np), np, red);
    if (threadIdx.x != 0) return;
    st->rz_prev[0] = st->rz[0];
    st->rz_prev[1] = st->rz[1];
    ctrl_check_op(st, prm, hist ? hist + size_t(r) * hist_stride : nullptr, s);
}

// beta from the new rz / rho partials (and the COCR rho breakdown stop)
template <typename T>
__global__ __launch_bounds__(kCtrlThreads) void k_ctrl_beta_blk(const T* __restrict__ part, int np,
                                                               SolverState* st, SolveParams prm)
{
    const int r = blockIdx.x;
    st += r;
    if (st->status) return;
    __shared__ T red[kCtrlThreads / 64];
    const T v = sum_partials(part + size_t(r) * size_t(np), np, red);
    if (threadIdx.x == 0) ctrl_beta_op(st, prm, v, from_pair<T>(st->rz_prev));
}

}  // namespace ssa
