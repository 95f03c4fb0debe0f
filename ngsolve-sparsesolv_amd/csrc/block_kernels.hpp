// block_kernels.hpp -- vector and control kernels for a block of W
// right-hand sides of one system (interleaved rows: element (i, r) at
// v[i*W + r]; the SpMV and the sweeps take W as a template in kernels.hpp).
//
// The kernels that reduce (dots, norms, the residual updates) sum, for
// right-hand side r in workgroup g, the contributions of rows g*kBlock + t
// (stepping by G*kBlock) in thread t, in the same order, with block_sum's
// tree -- as workgroup g of the one-vector kernel does -- so a block solve
// reproduces the one-vector solves bit for bit (tile_accumulate,
// tile_finish).  Element-wise
// kernels take the flat, coalesced mapping (thread e, element e; G*kBlock is
// a multiple of W, so a thread keeps one right-hand side).  The control steps
// run as W one-workgroup kernels between producer and consumer: a fused
// step would have every consuming workgroup re-sum W sets of partials.
// Partials of right-hand side r occupy [r*np, (r+1)*np) for a producer of
// np workgroups.
#pragma once

#include "kernels.hpp"

namespace ssa {

// Reductions over a block keep the one-vector kernel's tree: in workgroup g,
// thread t owns rows g*kBlock + t (stepping by G*kBlock) and adds their
// contributions in that order, then block_sum combines the threads.  The
// memory side stays coalesced: a tile of kBlock rows is processed element by
// element (consecutive threads, consecutive elements), CP columns at a time.
// In pass h a thread keeps one column, h*CP + (t & (CP-1)), and the rows
// j*(kBlock/CP) + t/CP (j < CP) of every tile; it accumulates each
// (row, column) pair in its own register, acc[h*CP + j], tile after tile --
// the owner's sequence of additions, done by another thread.  tile_finish
// hands every sum to its owning thread through LDS once at the end (a
// kBlock x (CP+1) tile per pass, one pad column against bank conflicts)
// instead of once per tile.
template <typename V, int W>
constexpr int pass_cols()
{
    constexpr int lim = int(160 / sizeof(V)) - 1;
    return W <= lim ? W : lim >= 8 ? 8 : lim >= 4 ? 4 : 1;
}

// columns per exchange in tile_finish (an LDS tile of kBlock x (FK+1)
// values, at most 10 KB for FP64 and 20 KB for complex<double>)
template <typename V, int W>
constexpr int finish_cols()
{
    constexpr int CP = pass_cols<V, W>();
    constexpr int lim = sizeof(V) >= 16 ? 2 : sizeof(V) >= 8 ? 4 : 8;
    return CP < lim ? CP : lim;
}
template <typename V, int W>
constexpr int tile_words() { return kBlock * (finish_cols<V, W>() + 1); }

// One tile: rows [base, base + rows) (rows <= kBlock).  fn(e, h) does the
// element-wise work of element e (index into the interleaved vectors; its
// column is h*CP + (threadIdx.x & (CP-1))) and returns the contribution.
template <typename V, int W, typename Fn>
__device__ inline void tile_accumulate(index_t base, index_t rows, V (&acc)[W], Fn&& fn)
{
    constexpr int CP = pass_cols<V, W>();
    const int t = threadIdx.x;
#pragma unroll
    for (int h = 0; h < W / CP; ++h) {
        const int c = h * CP + (t & (CP - 1));
#pragma unroll
        for (int j = 0; j < CP; ++j) {
            const int lr = (j * kBlock + t) / CP;
            if (lr < rows) acc[h * CP + j] = add(acc[h * CP + j], fn(size_t(base + lr) * W + c, h));
        }
    }
}

// s[c] of thread t <- the sum of row t, column c (tile_accumulate's acc of
// the thread that kept it), FK columns per exchange; `tile` holds
// tile_words<V, W>() values
template <typename V, int W>
__device__ inline void tile_finish(const V (&acc)[W], V* tile, V (&s)[W])
{
    constexpr int CP = pass_cols<V, W>();
    constexpr int FK = finish_cols<V, W>();
    const int t = threadIdx.x;
    const int ct = t & (CP - 1);
#pragma unroll
    for (int h = 0; h < W / CP; ++h) {
#pragma unroll
        for (int k0 = 0; k0 < CP; k0 += FK) {
            if (ct >= k0 && ct < k0 + FK) {
#pragma unroll
                for (int j = 0; j < CP; ++j) {
                    const int lr = (j * kBlock + t) / CP;
                    tile[lr * (FK + 1) + (ct - k0)] = acc[h * CP + j];
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < FK; ++k) s[h * CP + k0 + k] = tile[t * (FK + 1) + k];
            __syncthreads();
        }
    }
}

// the column of thread t in pass h of a tile with CP columns per pass
template <typename V, int W>
__device__ inline int pass_column(int h)
{
    constexpr int CP = pass_cols<V, W>();
    return h * CP + int(threadIdx.x & (CP - 1));
}

// partial sums of conj?(a_i) b_i, accumulated in T (operands converted).
// Complex at W = 16: held to 128 VGPRs (4 waves per SIMD, so the kMaxGrid
// workgroups are resident at once; 130 VGPRs gave 3 and a second, quarter
// full round): eddy78 (q, qt) dot 134 -> 120 us
template <typename T, typename TA, typename TB, int W>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(is_cplx<T>::value && W >= 16 ? 4 : 1))) void k_dot_blk(index_t n, const TA* __restrict__ a,
                                                   const TB* __restrict__ b, int conj,
                                                   T* __restrict__ partials)
{
    __shared__ T red[kWavesPerBlock * W];
    __shared__ T tile[tile_words<T, W>()];
    T acc[W], d[W];
#pragma unroll
    for (int r = 0; r < W; ++r) acc[r] = zero<T>();
    for (index_t base = index_t(blockIdx.x) * kBlock; base < n; base += index_t(gridDim.x) * kBlock) {
        const index_t rows = n - base < kBlock ? n - base : kBlock;
        tile_accumulate<T, W>(base, rows, acc, [&](size_t e, int) {
            return mul(conj_if(cvt<T>(a[e]), conj != 0), cvt<T>(b[e]));
        });
    }
    tile_finish<T, W>(acc, tile, d);
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
    __shared__ double tile[tile_words<double, W>()];
    double aa[W], ab[W], sa[W], sb[W];
#pragma unroll
    for (int r = 0; r < W; ++r) aa[r] = ab[r] = 0.0;
    for (index_t base = index_t(blockIdx.x) * kBlock; base < n; base += index_t(gridDim.x) * kBlock) {
        const index_t rows = n - base < kBlock ? n - base : kBlock;
        tile_accumulate<double, W>(base, rows, aa, [&](size_t e, int) { return norm2(a[e]); });
        if (b) tile_accumulate<double, W>(base, rows, ab, [&](size_t e, int) { return norm2(b[e]); });
    }
    tile_finish<double, W>(aa, tile, sa);
    if (b) tile_finish<double, W>(ab, tile, sb);
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
    constexpr int CP = pass_cols<double, W>();
    __shared__ double tile[tile_words<double, W>()];
    // the column a thread updates in each pass, its alpha and whether it runs
    T alpha[W / CP];
    bool run[W / CP];
#pragma unroll
    for (int h = 0; h < W / CP; ++h) {
        const int c = pass_column<double, W>(h);
        run[h] = st[c].status == 0;
        alpha[h] = from_pair<T>(st[c].alpha);
    }
    double acc[W], s[W];
#pragma unroll
    for (int q = 0; q < W; ++q) acc[q] = 0.0;
    for (index_t base = index_t(blockIdx.x) * kBlock; base < n; base += index_t(gridDim.x) * kBlock) {
        const index_t rows = n - base < kBlock ? n - base : kBlock;
        tile_accumulate<double, W>(base, rows, acc, [&](size_t e, int h) {
            if (!run[h]) return 0.0;
            x[e] = fma_(alpha[h], p[e], x[e]);
            const T ri = fms_(alpha[h], Ap[e], r[e]);
            r[e] = ri;
            if (r_low) r_low[e] = cvt<P>(ri);
            return norm2(ri);
        });
    }
    tile_finish<double, W>(acc, tile, s);
    const double t = block_sum_multi<double, W>(s, red);
    if (threadIdx.x < W) partials[size_t(threadIdx.x) * gridDim.x + blockIdx.x] = t;
}

// COCR: the previous iteration's direction update p = rt + beta p (beta
// from the state; 0 in a pass's first iteration, where p holds rt), then
// x += alpha p; r -= alpha q; rt -= alpha qt; partial |r|^2 (qt and rt may
// be in the preconditioner's precision, rt computed wide, stored narrow).
// p is read here anyway (x += alpha p), so the p, q update after the SpMV
// leaves p alone: the same operations on the same operands as the
// one-vector loop (p updated there, read here), two vectors fewer through
// HBM.
// best <- x (before its update) for the right-hand sides whose last check
// improved, as the one-vector k_cocr_update does.
template <typename T, typename TQ, int W, typename TR = T>
__global__ __launch_bounds__(kBlock) void k_cocr_update_blk(index_t n, T* __restrict__ x,
                                                           T* __restrict__ r, TR* __restrict__ rt,
                                                           T* __restrict__ p,
                                                           const T* __restrict__ q,
                                                           const TQ* __restrict__ qt,
                                                           double* __restrict__ partials,
                                                           const SolverState* __restrict__ st,
                                                           T* __restrict__ best)
{
    if (all_stopped<W>(st)) return;
    __shared__ double red[kWavesPerBlock * W];
    constexpr int CP = pass_cols<double, W>();
    __shared__ double tile[tile_words<double, W>()];
    T alpha[W / CP], beta[W / CP];
    bool run[W / CP], copy[W / CP];
#pragma unroll
    for (int h = 0; h < W / CP; ++h) {
        const int c = pass_column<double, W>(h);
        run[h] = st[c].status == 0;
        copy[h] = best && st[c].improved;
        alpha[h] = from_pair<T>(st[c].alpha);
        beta[h] = from_pair<T>(st[c].beta);
    }
    double acc[W], s[W];
#pragma unroll
    for (int k = 0; k < W; ++k) acc[k] = 0.0;
    for (index_t base = index_t(blockIdx.x) * kBlock; base < n; base += index_t(gridDim.x) * kBlock) {
        const index_t rows = n - base < kBlock ? n - base : kBlock;
        tile_accumulate<double, W>(base, rows, acc, [&](size_t e, int h) {
            if (!run[h]) return 0.0;
            const T rte = cvt<T>(rt[e]);
            const T pe = fma_(beta[h], load_nt(p + e), rte);
            store_nt(p + e, pe);
            const T xe = load_nt(x + e);
            if (copy[h]) store_nt(best + e, xe);
            store_nt(x + e, fma_(alpha[h], pe, xe));
            const T ri = fms_(alpha[h], load_nt(q + e), load_nt(r + e));
            store_nt(r + e, ri);
            rt[e] = cvt<TR>(fms_(alpha[h], cvt<T>(load_nt(qt + e)), rte));
            return norm2(ri);
        });
    }
    tile_finish<double, W>(acc, tile, s);
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

// COCR: q = t + beta q (q also in the preconditioner's precision when q_low
// is given).  p = rt + beta p and the best-x copy are left to the next
// iteration's k_cocr_update_blk.
template <typename T, typename P, int W>
__global__ __launch_bounds__(kBlock) void k_cocr_q_blk(size_t total, const T* __restrict__ t,
                                                      T* __restrict__ q,
                                                      const SolverState* __restrict__ st,
                                                      P* __restrict__ q_low)
{
    const int k = int(threadIdx.x) & (W - 1);
    if (st[k].status != 0) return;
    const T beta = from_pair<T>(st[k].beta);
    for (size_t e = size_t(blockIdx.x) * kBlock + threadIdx.x; e < total;
         e += size_t(gridDim.x) * kBlock) {
        const T qi = fma_(beta, load_nt(q + e), load_nt(t + e));
        store_nt(q + e, qi);
        if (q_low) q_low[e] = cvt<P>(qi);
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
