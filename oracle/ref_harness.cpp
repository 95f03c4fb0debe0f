// ref_harness.cpp -- exports the entry points of oracle.h, implemented by
// the reference header library itself (include/sparsesolv of
// ksugahar/ngsolve-sparsesolv, compiled from where it lies, never copied).
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/,
// which is git-ignored; it is used to pin the C restatement in oracle.c
// and as the CPU baseline of bench.py.  The product never loads it.
//
// The reference keeps the factor in private members of ICPreconditioner;
// the harness reads them by compiling the reference headers with private
// access opened, after every standard header they use is already included.

#include <algorithm>
#include <atomic>
#include <cassert>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <iostream>
#include <limits>
#include <map>
#include <memory>
#include <numeric>
#include <queue>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#define private public
#define protected public
#include <sparsesolv/sparsesolv.hpp>
#undef private
#undef protected

#include "oracle.h"

using sparsesolv::index_t;
using cplx = std::complex<double>;

namespace {

sparsesolv::SolverConfig to_config(const orc_config* c)
{
    sparsesolv::SolverConfig s;
    s.tolerance = c->tolerance;
    s.max_iterations = c->max_iterations;
    s.abs_tolerance = c->abs_tolerance;
    s.norm_type = c->norm_type == 0 ? sparsesolv::NormType::RHS
                : c->norm_type == 1 ? sparsesolv::NormType::InitialResidual
                                    : sparsesolv::NormType::Custom;
    s.custom_norm = c->custom_norm;
    s.shift_parameter = c->shift_parameter;
    s.diagonal_scaling = c->diagonal_scaling != 0;
    s.auto_shift = c->auto_shift != 0;
    s.shift_increment = c->shift_increment;
    s.max_shift_value = c->max_shift_value;
    s.min_diagonal_threshold = c->min_diagonal_threshold;
    s.zero_diagonal_replacement = c->zero_diagonal_replacement;
    s.max_shift_trials = c->max_shift_trials;
    s.divergence_check = c->divergence_check ? sparsesolv::DivergenceCheck::StagnationCount
                                             : sparsesolv::DivergenceCheck::None;
    s.divergence_threshold = c->divergence_threshold;
    s.divergence_count = c->divergence_count;
    s.save_best_result = c->save_best_result != 0;
    s.save_residual_history = c->save_residual_history != 0;
    s.conjugate = c->conjugate != 0;
    s.use_abmc = c->use_abmc != 0;
    s.abmc_block_size = c->abmc_block_size;
    s.abmc_num_colors = c->abmc_num_colors;
    s.abmc_reorder_spmv = c->abmc_reorder_spmv != 0;
    s.abmc_use_rcm = c->abmc_use_rcm != 0;
    return s;
}

void put_err(char* err, int errlen, const char* msg)
{
    if (err && errlen > 0) {
        std::strncpy(err, msg, static_cast<size_t>(errlen) - 1);
        err[errlen - 1] = 0;
    }
}

struct IcAny {
    int dtype;
    int n;
    std::unique_ptr<sparsesolv::ICPreconditioner<double>> d;
    std::unique_ptr<sparsesolv::ICPreconditioner<cplx>> z;
};

template <typename T>
void fill_result(const sparsesolv::SolverResult& r, orc_result* res, double* hist,
                 double shift)
{
    res->converged = r.converged ? 1 : 0;
    res->iterations = r.iterations;
    res->final_residual = r.final_residual;
    res->history_len = static_cast<int>(r.residual_history.size());
    res->actual_shift = shift;
    if (hist)
        std::copy(r.residual_history.begin(), r.residual_history.end(), hist);
}

template <typename T>
int solve_impl(int method, int n, const int* rp, const int* ci, const void* values,
               const void* b, void* x, const orc_config* cfg, orc_result* res,
               double* hist, char* err, int errlen)
{
    try {
        sparsesolv::SparseMatrixView<T> A(n, n, rp, ci, static_cast<const T*>(values));
        const T* bb = static_cast<const T*>(b);
        T* xx = static_cast<T*>(x);
        sparsesolv::SolverConfig c = to_config(cfg);
        sparsesolv::SolverResult r;
        double shift = 0.0;
        if (method == 0) {
            // solve_iccg rebuilds the preconditioner internally; replicate
            // its shift report by one extra setup on the same config.
            r = sparsesolv::solve_iccg<T>(A, bb, xx, n, c);
            sparsesolv::ICPreconditioner<T> pc(c.shift_parameter);
            pc.set_config(c);
            pc.setup(A);
            shift = pc.actual_shift();
        } else if (method == 1) {
            r = sparsesolv::solve_cocr<T>(A, bb, xx, n, c);
            sparsesolv::ICPreconditioner<T> pc(c.shift_parameter);
            pc.set_config(c);
            pc.setup(A);
            shift = pc.actual_shift();
        } else if (method == 2) {
            sparsesolv::CGSolver<T> s;
            s.set_config(c);
            r = s.solve(A, bb, xx, n, nullptr);
        } else if (method == 3) {
            sparsesolv::COCRSolver<T> s;
            s.set_config(c);
            r = s.solve(A, bb, xx, n, nullptr);
        } else if (method == 4 || method == 5) {
            // CGSolver / COCRSolver with an SGSPreconditioner set up on A
            sparsesolv::SGSPreconditioner<T> pc;
            pc.setup(A);
            if (method == 4) {
                sparsesolv::CGSolver<T> s;
                s.set_config(c);
                r = s.solve(A, bb, xx, n, &pc);
            } else {
                sparsesolv::COCRSolver<T> s;
                s.set_config(c);
                r = s.solve(A, bb, xx, n, &pc);
            }
        } else {
            r = sparsesolv::solve_sgsmrtr<T>(A, bb, xx, n, c);
        }
        fill_result<T>(r, res, hist, shift);
        return 0;
    } catch (const std::exception& e) {
        put_err(err, errlen, e.what());
        return -1;
    }
}

// SGSPreconditioner::setup + apply
template <typename T>
void sgs_apply_impl(int n, const int* rp, const int* ci, const void* values, const void* x, void* y)
{
    sparsesolv::SparseMatrixView<T> A(n, n, rp, ci, static_cast<const T*>(values));
    sparsesolv::SGSPreconditioner<T> pc;
    pc.setup(A);
    pc.apply(static_cast<const T*>(x), static_cast<T*>(y), n);
}

// COCRSolver::solve(A, b, x, n, &pc) with pc set up on another matrix M
template <typename T>
int solve_pre_impl(int n, const int* rp, const int* ci, const void* values, const int* mrp,
                   const int* mci, const void* mvalues, const void* b, void* x,
                   const orc_config* cfg, orc_result* res, double* hist, char* err, int errlen)
{
    try {
        sparsesolv::SparseMatrixView<T> A(n, n, rp, ci, static_cast<const T*>(values));
        sparsesolv::SparseMatrixView<T> M(n, n, mrp, mci, static_cast<const T*>(mvalues));
        sparsesolv::SolverConfig c = to_config(cfg);
        sparsesolv::ICPreconditioner<T> pc(c.shift_parameter);
        pc.set_config(c);
        pc.setup(M);
        sparsesolv::COCRSolver<T> s;
        s.set_config(c);
        const sparsesolv::SolverResult r =
            s.solve(A, static_cast<const T*>(b), static_cast<T*>(x), n, &pc);
        fill_result<T>(r, res, hist, pc.actual_shift());
        return 0;
    } catch (const std::exception& e) {
        put_err(err, errlen, e.what());
        return -1;
    }
}

template <typename T>
double time_impl(int method, int n, const int* rp, const int* ci, const void* values,
                 const void* b, const orc_config* cfg, int iters, double* setup_s)
{
    sparsesolv::SparseMatrixView<T> A(n, n, rp, ci, static_cast<const T*>(values));
    sparsesolv::SolverConfig c = to_config(cfg);
    c.tolerance = 0.0;
    c.abs_tolerance = 0.0;
    c.max_iterations = iters;
    c.save_residual_history = false;
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    sparsesolv::ICPreconditioner<T> pc(c.shift_parameter);
    pc.set_config(c);
    pc.setup(A);
    auto t1 = clk::now();
    if (setup_s) *setup_s = std::chrono::duration<double>(t1 - t0).count();
    std::vector<T> x(static_cast<size_t>(n), T(0));
    auto t2 = clk::now();
    if (method == 0) {
        sparsesolv::CGSolver<T> s;
        s.set_config(c);
        s.solve(A, static_cast<const T*>(b), x.data(), n, &pc);
    } else {
        sparsesolv::COCRSolver<T> s;
        s.set_config(c);
        s.solve(A, static_cast<const T*>(b), x.data(), n, &pc);
    }
    auto t3 = clk::now();
    return std::chrono::duration<double>(t3 - t2).count();
}

template <typename P>
static int count_of(const P& p, int n, int which)
{
    switch (which) {
    case 0: return n;
    case 1: return static_cast<int>(p.L_.values.size());
    case 2: return p.config_.use_abmc ? p.abmc_schedule_.num_colors() : 0;
    case 3: return p.config_.use_abmc ? p.abmc_schedule_.num_blocks() : 0;
    default: return -1;
    }
}

template <typename P>
static void copy_int_of(const P& p, int which, int* out)
{
    const std::vector<index_t>* v = nullptr;
    switch (which) {
    case 0: v = &p.L_.row_ptr; break;
    case 1: v = &p.L_.col_idx; break;
    case 2: v = &p.Lt_.row_ptr; break;
    case 3: v = &p.Lt_.col_idx; break;
    case 4: v = &p.abmc_schedule_.ordering; break;
    case 5: v = &p.abmc_schedule_.color_offsets; break;
    case 6: v = &p.abmc_schedule_.block_offsets; break;
    default: return;
    }
    std::copy(v->begin(), v->end(), out);
}

template <typename T, typename P>
static void copy_val_of(const P& p, int which, void* out)
{
    T* o = static_cast<T*>(out);
    switch (which) {
    case 0: std::copy(p.L_.values.begin(), p.L_.values.end(), o); break;
    case 1: std::copy(p.Lt_.values.begin(), p.Lt_.values.end(), o); break;
    case 2: std::copy(p.inv_diag_.begin(), p.inv_diag_.end(), o); break;
    case 3: std::copy(p.scaling_.begin(), p.scaling_.end(), static_cast<double*>(out)); break;
    default: break;
    }
}

}  // namespace

extern "C" {

void orc_config_default(orc_config* o)
{
    sparsesolv::SolverConfig s;
    o->tolerance = s.tolerance;
    o->max_iterations = s.max_iterations;
    o->abs_tolerance = s.abs_tolerance;
    o->norm_type = 0;
    o->custom_norm = s.custom_norm;
    o->shift_parameter = s.shift_parameter;
    o->diagonal_scaling = s.diagonal_scaling;
    o->auto_shift = s.auto_shift;
    o->shift_increment = s.shift_increment;
    o->max_shift_value = s.max_shift_value;
    o->min_diagonal_threshold = s.min_diagonal_threshold;
    o->zero_diagonal_replacement = s.zero_diagonal_replacement;
    o->max_shift_trials = s.max_shift_trials;
    o->divergence_check = 0;
    o->divergence_threshold = s.divergence_threshold;
    o->divergence_count = s.divergence_count;
    o->save_best_result = s.save_best_result;
    o->save_residual_history = s.save_residual_history;
    o->conjugate = s.conjugate;
    o->use_abmc = s.use_abmc;
    o->abmc_block_size = s.abmc_block_size;
    o->abmc_num_colors = s.abmc_num_colors;
    o->abmc_reorder_spmv = s.abmc_reorder_spmv;
    o->abmc_use_rcm = s.abmc_use_rcm;
}

void* orc_abmc_build(int n, const int* row_ptr, const int* col_idx, int block_size,
                     int target_colors)
{
    auto* s = new sparsesolv::ABMCSchedule();
    s->build(row_ptr, col_idx, n, block_size, target_colors);
    return s;
}

int orc_abmc_count(const void* h, int which)
{
    auto* s = static_cast<const sparsesolv::ABMCSchedule*>(h);
    return which == 0 ? s->num_colors() : s->num_blocks();
}

void orc_abmc_copy(const void* h, int which, int* out)
{
    auto* s = static_cast<const sparsesolv::ABMCSchedule*>(h);
    const std::vector<index_t>* v = nullptr;
    switch (which) {
    case 0: v = &s->color_offsets; break;
    case 1: v = &s->color_blocks; break;
    case 2: v = &s->block_offsets; break;
    case 3: v = &s->ordering; break;
    case 4: v = &s->reverse_ordering; break;
    default: return;
    }
    std::copy(v->begin(), v->end(), out);
}

void orc_abmc_free(void* h) { delete static_cast<sparsesolv::ABMCSchedule*>(h); }

void* orc_ic_setup(int dtype, int n, const int* row_ptr, const int* col_idx,
                   const void* values, const orc_config* cfg, char* err, int errlen)
{
    auto* h = new IcAny();
    h->dtype = dtype;
    h->n = n;
    try {
        sparsesolv::SolverConfig c = to_config(cfg);
        if (dtype == 0) {
            sparsesolv::SparseMatrixView<double> A(n, n, row_ptr, col_idx,
                                                   static_cast<const double*>(values));
            h->d.reset(new sparsesolv::ICPreconditioner<double>(c.shift_parameter));
            h->d->set_config(c);
            h->d->setup(A);
        } else {
            sparsesolv::SparseMatrixView<cplx> A(n, n, row_ptr, col_idx,
                                                 static_cast<const cplx*>(values));
            h->z.reset(new sparsesolv::ICPreconditioner<cplx>(c.shift_parameter));
            h->z->set_config(c);
            h->z->setup(A);
        }
    } catch (const std::exception& e) {
        put_err(err, errlen, e.what());
        delete h;
        return nullptr;
    }
    return h;
}

int orc_ic_count(const void* hv, int which)
{
    auto* h = static_cast<const IcAny*>(hv);
    return h->dtype == 0 ? count_of(*h->d, h->n, which) : count_of(*h->z, h->n, which);
}

double orc_ic_actual_shift(const void* hv)
{
    auto* h = static_cast<const IcAny*>(hv);
    return h->dtype == 0 ? h->d->actual_shift() : h->z->actual_shift();
}

void orc_ic_copy_int(const void* hv, int which, int* out)
{
    auto* h = static_cast<const IcAny*>(hv);
    if (h->dtype == 0) copy_int_of(*h->d, which, out); else copy_int_of(*h->z, which, out);
}

void orc_ic_copy_val(const void* hv, int which, void* out)
{
    auto* h = static_cast<const IcAny*>(hv);
    if (h->dtype == 0) copy_val_of<double>(*h->d, which, out);
    else copy_val_of<cplx>(*h->z, which, out);
}

void orc_ic_apply(const void* hv, const void* x, void* y)
{
    auto* h = static_cast<const IcAny*>(hv);
    if (h->dtype == 0)
        h->d->apply(static_cast<const double*>(x), static_cast<double*>(y), h->n);
    else
        h->z->apply(static_cast<const cplx*>(x), static_cast<cplx*>(y), h->n);
}

void orc_ic_apply_reordered(const void* hv, const void* x, void* y)
{
    auto* h = static_cast<const IcAny*>(hv);
    if (h->dtype == 0)
        h->d->apply_in_reordered_space(static_cast<const double*>(x),
                                       static_cast<double*>(y), h->n);
    else
        h->z->apply_in_reordered_space(static_cast<const cplx*>(x),
                                       static_cast<cplx*>(y), h->n);
}

void orc_ic_free(void* hv) { delete static_cast<IcAny*>(hv); }

int orc_solve_iccg(int dtype, int n, const int* rp, const int* ci, const void* values,
                   const void* b, void* x, const orc_config* cfg, orc_result* res,
                   double* hist, char* err, int errlen)
{
    return dtype == 0 ? solve_impl<double>(0, n, rp, ci, values, b, x, cfg, res, hist, err, errlen)
                      : solve_impl<cplx>(0, n, rp, ci, values, b, x, cfg, res, hist, err, errlen);
}

int orc_solve_cocr(int dtype, int n, const int* rp, const int* ci, const void* values,
                   const void* b, void* x, const orc_config* cfg, orc_result* res,
                   double* hist, char* err, int errlen)
{
    return dtype == 0 ? solve_impl<double>(1, n, rp, ci, values, b, x, cfg, res, hist, err, errlen)
                      : solve_impl<cplx>(1, n, rp, ci, values, b, x, cfg, res, hist, err, errlen);
}

int orc_solve_cocr_pre(int dtype, int n, const int* rp, const int* ci, const void* values,
                       const int* mrp, const int* mci, const void* mvalues, const void* b,
                       void* x, const orc_config* cfg, orc_result* res, double* hist,
                       char* err, int errlen)
{
    return dtype == 0 ? solve_pre_impl<double>(n, rp, ci, values, mrp, mci, mvalues, b, x, cfg,
                                               res, hist, err, errlen)
                      : solve_pre_impl<cplx>(n, rp, ci, values, mrp, mci, mvalues, b, x, cfg,
                                             res, hist, err, errlen);
}

int orc_solve_plain(int dtype, int method_cocr, int n, const int* rp, const int* ci,
                    const void* values, const void* b, void* x, const orc_config* cfg,
                    orc_result* res, double* hist, char* err, int errlen)
{
    const int m = method_cocr ? 3 : 2;
    return dtype == 0 ? solve_impl<double>(m, n, rp, ci, values, b, x, cfg, res, hist, err, errlen)
                      : solve_impl<cplx>(m, n, rp, ci, values, b, x, cfg, res, hist, err, errlen);
}

void orc_sgs_apply(int dtype, int n, const int* rp, const int* ci, const void* values,
                   const void* x, void* y)
{
    if (dtype == 0) sgs_apply_impl<double>(n, rp, ci, values, x, y);
    else sgs_apply_impl<cplx>(n, rp, ci, values, x, y);
}

int orc_solve_sgs(int dtype, int method_cocr, int n, const int* rp, const int* ci,
                  const void* values, const void* b, void* x, const orc_config* cfg,
                  orc_result* res, double* hist, char* err, int errlen)
{
    const int m = method_cocr ? 5 : 4;
    return dtype == 0 ? solve_impl<double>(m, n, rp, ci, values, b, x, cfg, res, hist, err, errlen)
                      : solve_impl<cplx>(m, n, rp, ci, values, b, x, cfg, res, hist, err, errlen);
}

int orc_solve_sgsmrtr(int dtype, int n, const int* rp, const int* ci, const void* values,
                      const void* b, void* x, const orc_config* cfg, orc_result* res,
                      double* hist, char* err, int errlen)
{
    return dtype == 0 ? solve_impl<double>(6, n, rp, ci, values, b, x, cfg, res, hist, err, errlen)
                      : solve_impl<cplx>(6, n, rp, ci, values, b, x, cfg, res, hist, err, errlen);
}

double orc_time_iterations(int dtype, int method, int n, const int* rp, const int* ci,
                           const void* values, const void* b, const orc_config* cfg,
                           int iters, double* setup_s)
{
    try {
        return dtype == 0 ? time_impl<double>(method, n, rp, ci, values, b, cfg, iters, setup_s)
                          : time_impl<cplx>(method, n, rp, ci, values, b, cfg, iters, setup_s);
    } catch (const std::exception&) {
        return -1.0;
    }
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// threads of the following parallel regions (OpenMP build; no effect otherwise)
void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

}  // extern "C"
