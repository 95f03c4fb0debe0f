// mrtr_kernels.hpp -- the vector and control kernels of SGS-MRTR
// (SGSMRTRSolver, solvers/sgs_mrtr_solver.hpp:53-365) on one system.
//
// The iteration runs in the diagonally scaled space of D A D, D =
// diag(1/sqrt|a_ii|), with L = lower(D A D) (diagonal included).  Its
// triangular solves are the level-schedule sweeps of the IC path:
//   forward  w = L^-1 v             : w_i = (v_i - sum_{k<i} L_ik w_k) / L_ii
//   backward u = L^-T rd            : the sweep computes
//            u_i = s_i - (1/L_ii) sum_{k>i} L_ki u_k with s = rd / L_ii,
//            which is (rd_i - sum L_ki u_k) / L_ii.
// The scalars (zeta, eta, nu and the p coefficient) live in SolverState:
// alpha = zeta, rz_prev = zeta_old, beta = eta, rz = nu, denom = the p
// coefficient eta zeta_old / zeta.  Every kernel returns at once after a
// stop, so captured iterations can be replayed past it.
#pragma once

#include "kernels.hpp"

namespace ssa {

// prepare_iteration (sgs_mrtr_solver.hpp:101-116): r = D (b - A x) into r,
// b2 = D b over b, x2 = D^-1 x over x; partials of |r|^2 and |b2|^2
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mrtr_prep(index_t n, const T* __restrict__ d,
                                                     const T* __restrict__ id, T* __restrict__ b,
                                                     const T* __restrict__ Ax, T* __restrict__ x,
                                                     T* __restrict__ r, double* __restrict__ pr,
                                                     double* __restrict__ pb)
{
    __shared__ double red[kWavesPerBlock];
    double sr = 0.0, sb = 0.0;
    for (index_t i = index_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += index_t(gridDim.x) * kBlock) {
        const T di = d[i], bi = b[i];
        const T ri = mul(di, sub(bi, Ax[i]));
        const T b2 = mul(di, bi);
        r[i] = ri;
        b[i] = b2;
        x[i] = mul(id[i], x[i]);
        sr += norm2(ri);
        sb += norm2(b2);
    }
    double t = block_sum(sr, red);
    if (threadIdx.x == 0) pr[blockIdx.x] = t;
    t = block_sum(sb, red);
    if (threadIdx.x == 0) pb[blockIdx.x] = t;
}

// the normaliser's partials: |L b2|^2 by rows of L (CSR, one row a thread;
// the product is not kept) (sgs_mrtr_solver.hpp:119-121)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mrtr_lnorm(index_t n, const index_t* __restrict__ rp,
                                                      const index_t* __restrict__ ci,
                                                      const T* __restrict__ v,
                                                      const T* __restrict__ b2,
                                                      double* __restrict__ part)
{
    __shared__ double red[kWavesPerBlock];
    double s = 0.0;
    for (index_t i = index_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += index_t(gridDim.x) * kBlock) {
        T acc = zero<T>();
        for (index_t k = rp[i]; k < rp[i + 1]; ++k) acc = fma_(v[k], b2[ci[k]], acc);
        s += norm2(acc);
    }
    const double t = block_sum(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// y = -rd; s = rd / L_ii (the backward sweep's input); partials of |rd|^2
// (sgs_mrtr_solver.hpp:141-146)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mrtr_start(index_t n, const T* __restrict__ rd,
                                                      const T* __restrict__ invd, T* __restrict__ y,
                                                      T* __restrict__ s, double* __restrict__ part)
{
    __shared__ double red[kWavesPerBlock];
    double sum = 0.0;
    for (index_t i = index_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += index_t(gridDim.x) * kBlock) {
        const T v = rd[i];
        y[i] = sub(zero<T>(), v);
        s[i] = mul(v, invd[i]);
        sum += norm2(v);
    }
    const double t = block_sum(sum, red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// prepare_iteration's scalars and do_iterate's early exit (one workgroup):
// normaliser |L b2| (1 below 1e-15), the first history entry |r| / it
// (sgs_mrtr_solver.hpp:119-138), and converged with no iteration when
// |r| / |b2| < tolerance / 10 (:152-160; |b2| below 1e-30 counts as 1).  A
// budget of no iterations stops with the final residual |rd| (:242-244).
static __global__ __launch_bounds__(kCtrlThreads) void k_mrtr_init(const double* __restrict__ pr,
                                                                   const double* __restrict__ pb,
                                                                   const double* __restrict__ pl,
                                                                   const double* __restrict__ pd,
                                                                   int np, SolverState* st,
                                                                   SolveParams prm,
                                                                   double* __restrict__ hist)
{
    __shared__ double red[kCtrlThreads / 64];
    const double sr = sum_partials(pr, np, red);
    const double sb = sum_partials(pb, np, red);
    const double sl = sum_partials(pl, np, red);
    const double sd = sum_partials(pd, np, red);
    if (threadIdx.x != 0) return;
    const double nl = sqrt(sl), norm_r = sqrt(sr);
    double norm_b2 = sqrt(sb);
    if (norm_b2 < 1e-30) norm_b2 = 1.0;
    st->normalizer = nl > 1e-15 ? nl : 1.0;
    st->norm_r = norm_r;
    st->best_residual = DBL_MAX;
    st->bad_count = 0;
    st->improved = 0;
    st->iter = 0;
    st->hist_len = 0;
    st->status = kRunning;
    st->converged = 0;
    st->tol = prm.tolerance;
    st->maxit = prm.max_iterations;
    if (prm.save_history) hist[st->hist_len++] = norm_r / st->normalizer;
    if (norm_r / norm_b2 < prm.tolerance * 0.1) {
        stop(st, kConverged, 1, 0, norm_r);
        return;
    }
    if (st->maxit <= 0) stop(st, kMaxIter, 0, 0, sqrt(sd));
}

// tmp = rd - u (the forward sweep's input, sgs_mrtr_solver.hpp:171-173)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mrtr_sub(index_t n, const T* __restrict__ rd,
                                                    const T* __restrict__ u, T* __restrict__ tmp,
                                                    const SolverState* __restrict__ st)
{
    if (st->status) return;
    for (index_t i = index_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += index_t(gridDim.x) * kBlock)
        tmp[i] = sub(rd[i], u[i]);
}

// Ard = w + u over w (:174-176); partials of (Ard, rd), (Ard, Ard),
// (Ard, y), conjugating the first operand when `conj` (IterativeSolver::
// dot_product, iterative_solver.hpp:303-322)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mrtr_dots(index_t n, T* __restrict__ w,
                                                     const T* __restrict__ u,
                                                     const T* __restrict__ rd,
                                                     const T* __restrict__ y, int conj,
                                                     T* __restrict__ p_r, T* __restrict__ p_a,
                                                     T* __restrict__ p_y,
                                                     const SolverState* __restrict__ st)
{
    if (st->status) return;
    __shared__ T red[kWavesPerBlock];
    T dr = zero<T>(), da = zero<T>(), dy = zero<T>();
    for (index_t i = index_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += index_t(gridDim.x) * kBlock) {
        const T a = add(w[i], u[i]);
        w[i] = a;
        const T ac = conj_if(a, conj != 0);
        dr = add(dr, mul(ac, rd[i]));
        da = add(da, mul(ac, a));
        dy = add(dy, mul(ac, y[i]));
    }
    T t = block_sum(dr, red);
    if (threadIdx.x == 0) p_r[blockIdx.x] = t;
    t = block_sum(da, red);
    if (threadIdx.x == 0) p_a[blockIdx.x] = t;
    t = block_sum(dy, red);
    if (threadIdx.x == 0) p_y[blockIdx.x] = t;
}

__device__ inline double neg(double a) { return -a; }
__device__ inline c128 neg(c128 a) { return c128{-a.x, -a.y}; }
__device__ inline bool nonneg_real(double a) { return a >= 0.0; }
__device__ inline bool nonneg_real(c128 a) { return a.x >= 0.0; }
// the real number v as a scalar of T's type (the first argument names it)
__device__ inline double as_real_scalar(double, double v) { return v; }
__device__ inline c128 as_real_scalar(c128, double v) { return c128{v, 0.0}; }

// zeta, eta, nu and the p coefficient (sgs_mrtr_solver.hpp:180-210), one
// workgroup.  The first iteration takes zeta = (Ar, r) / (Ar, Ar), eta = 0;
// later ones solve the 2x2 minimisation with the denominator
// nu (Ar, Ar) - (Ar, y)^2 held away from zero at +-1e-60 (by the sign of
// its real part).
template <typename T>
__global__ __launch_bounds__(kCtrlThreads) void k_mrtr_coef(const T* __restrict__ p_r,
                                                            const T* __restrict__ p_a,
                                                            const T* __restrict__ p_y, int np,
                                                            SolverState* st)
{
    if (st->status) return;
    __shared__ T red[kCtrlThreads / 64];
    const T ar_r = sum_partials(p_r, np, red);
    const T ar_ar = sum_partials(p_a, np, red);
    const T ar_y = sum_partials(p_y, np, red);
    if (threadIdx.x != 0) return;
    T zeta, eta, zeta_old;
    const T nu_old = from_pair<T>(st->rz);
    if (st->iter == 0) {
        zeta = ddiv(ar_r, ar_ar);
        zeta_old = zeta;
        eta = zero<T>();
    } else {
        zeta_old = from_pair<T>(st->rz_prev);
        T den = sub(mul(nu_old, ar_ar), mul(ar_y, ar_y));
        if (dabs(den) < 1e-60) den = as_real_scalar(den, nonneg_real(den) ? 1e-60 : -1e-60);
        const T inv = ddiv(as_real_scalar(zero<T>(), 1.0), den);
        zeta = mul(mul(nu_old, ar_r), inv);
        eta = mul(mul(neg(ar_y), ar_r), inv);
    }
    to_pair(mul(zeta, ar_r), st->rz);                       // nu
    to_pair(ddiv(mul(eta, zeta_old), zeta), st->denom);     // p coefficient
    to_pair(zeta, st->rz_prev);                             // zeta_old for the next iteration
    to_pair(zeta, st->alpha);
    to_pair(eta, st->beta);
}

// p = u + c p; x2 += zeta p; y = eta y + zeta Ard; rd -= y; s = rd / L_ii;
// partials of |rd|^2 (sgs_mrtr_solver.hpp:207-228)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mrtr_update(index_t n, const T* __restrict__ u,
                                                       const T* __restrict__ ard, T* __restrict__ p,
                                                       T* __restrict__ x2, T* __restrict__ y,
                                                       T* __restrict__ rd, T* __restrict__ s,
                                                       const T* __restrict__ invd,
                                                       double* __restrict__ part,
                                                       const SolverState* __restrict__ st)
{
    if (st->status) return;
    const T c = from_pair<T>(st->denom), zeta = from_pair<T>(st->alpha), eta = from_pair<T>(st->beta);
    __shared__ double red[kWavesPerBlock];
    double sum = 0.0;
    for (index_t i = index_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += index_t(gridDim.x) * kBlock) {
        const T pi = fma_(c, p[i], u[i]);
        p[i] = pi;
        x2[i] = fma_(zeta, pi, x2[i]);
        const T yi = fma_(eta, y[i], mul(zeta, ard[i]));
        y[i] = yi;
        const T ri = sub(rd[i], yi);
        rd[i] = ri;
        s[i] = mul(ri, invd[i]);
        sum += norm2(ri);
    }
    const double t = block_sum(sum, red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// check_convergence on |rd| (iterative_solver.hpp:217-249, with x_ = x2),
// the stagnation stop (sgs_mrtr_solver.hpp:235-239), then the loop index;
// at the iteration budget the solve stops with the last |rd| (:242-244).
// One workgroup.
static __global__ __launch_bounds__(kCtrlThreads) void k_mrtr_check(const double* __restrict__ part,
                                                                    int np, SolverState* st,
                                                                    SolveParams prm,
                                                                    double* __restrict__ hist)
{
    if (st->status) return;
    __shared__ double red[kCtrlThreads / 64];
    const double s = sum_partials(part, np, red);
    if (threadIdx.x != 0) return;
    ctrl_check_op(st, prm, hist, s);
    if (st->status) return;
    st->iter += 1;
    if (st->iter >= st->maxit) stop(st, kMaxIter, 0, st->maxit, st->norm_r);
}

// best <- x2 when the last check improved the best residual (also after
// the check that stopped the solve)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mrtr_best(index_t n, const T* __restrict__ x2,
                                                     T* __restrict__ best,
                                                     const SolverState* __restrict__ st)
{
    if (!st->improved) return;
    for (index_t i = index_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += index_t(gridDim.x) * kBlock)
        best[i] = x2[i];
}

// x = D x2 (sgs_mrtr_solver.hpp:278-282)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mrtr_unscale(index_t n, const T* __restrict__ d,
                                                        T* __restrict__ x)
{
    for (index_t i = index_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += index_t(gridDim.x) * kBlock)
        x[i] = mul(d[i], x[i]);
}

}  // namespace ssa
