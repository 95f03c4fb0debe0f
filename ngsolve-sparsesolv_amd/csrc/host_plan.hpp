// host_plan.hpp -- everything the host computes before the upload: the IC
// factor (in ABMC order), the operator in ABMC order, and the HBM layouts.
//
// build_plan is compiled once, by g++ without fused multiply-add
// (host_setup.cpp), so the factor is computed with the reference's own
// floating-point evaluation; the HIP translation unit only declares it.
#pragma once

#include <complex>
#include <vector>

#include "sparsesolv_amd/device_layout.hpp"
#include "sparsesolv_amd/ic_factor.hpp"

namespace ssa_host {

template <typename T, typename F>
struct HostPlan {
    sparsesolv_amd::IcFactorHost<T> fac;   // fac.A_perm: A in ABMC order
    sparsesolv_amd::SpmvLayout<T> spmv;
    sparsesolv_amd::SweepLayout<F> fwd, bwd;
    std::vector<F> invd;                   // inv_diag in the factor precision
    std::vector<double> scale;             // diagonal_scaling: 1/sqrt|a_ii| per device row
    // SGS-MRTR (factor kind kFactorSgsMrtr), in device order: L = lower(D A D)
    // as CSR (its product gives the normaliser) and D, 1/D
    sparsesolv_amd::CsrHost<T> lcsr;
    std::vector<T> split_d, split_id;
    double seconds = 0.0;                  // host setup wall time
};

// Throws std::runtime_error (reference messages) on failure.
template <typename T, typename F>
void build_plan(HostPlan<T, F>& plan, int n, const int* row_ptr, const int* col_idx,
                const T* values, const sparsesolv_amd::IcSettings& settings);

using cd = std::complex<double>;
using cf = std::complex<float>;

// Several independent systems prepared for one device handle.  Each system
// is planned on its own (build_plan), then the layouts are concatenated:
// system s owns the device rows [row0[s], row0[s] + nrows[s]) (padded to
// row0[s+1], a multiple of 256 rows, so a 4-slice SpMV workgroup never
// spans two systems).  With one system nothing is padded and the layouts
// are the system's own.  The caller's vectors hold the systems back to
// back in their own orderings (in_off).
template <typename T, typename F>
struct BatchPlan {
    int count = 0;
    std::vector<int> row0, nrows, in_off;
    std::vector<int> gather_in;     // device row -> caller index (-1: padding)
    std::vector<int> gather_out;    // caller index -> device row
    std::vector<int> slice_sys_f, slice_sys_b, walk_sys_f, walk_sys_b;
    sparsesolv_amd::SpmvLayout<T> spmv;
    sparsesolv_amd::SweepLayout<F> fwd, bwd;
    std::vector<F> invd;
    std::vector<double> scale;      // diagonal_scaling: each system's S in its rows
    // SGS-MRTR (one system only): see HostPlan
    sparsesolv_amd::CsrHost<T> lcsr;
    std::vector<T> split_d, split_id;
    // per system: setup facts kept after the layouts are combined
    std::vector<double> actual_shift, seconds;
    std::vector<int> shift_trials, num_colors, num_blocks, nnz, nnz_l;
    // the single system's full plan (count == 1), for the factor queries
    HostPlan<T, F> single;
    double wall_seconds = 0.0;
};

// Throws std::runtime_error (reference messages) on failure.
template <typename T, typename F>
void build_batch(BatchPlan<T, F>& plan, int count, const int* n, const int* const* row_ptr,
                 const int* const* col_idx, const T* const* values,
                 const sparsesolv_amd::IcSettings& settings);

extern template void build_plan<double, double>(HostPlan<double, double>&, int, const int*,
                                                const int*, const double*,
                                                const sparsesolv_amd::IcSettings&);
extern template void build_plan<double, float>(HostPlan<double, float>&, int, const int*,
                                               const int*, const double*,
                                               const sparsesolv_amd::IcSettings&);
extern template void build_plan<cd, cd>(HostPlan<cd, cd>&, int, const int*, const int*,
                                        const cd*, const sparsesolv_amd::IcSettings&);
extern template void build_plan<cd, cf>(HostPlan<cd, cf>&, int, const int*, const int*,
                                        const cd*, const sparsesolv_amd::IcSettings&);
#define SSA_DECLARE_BATCH(T, F)                                                                \
    extern template void build_batch<T, F>(BatchPlan<T, F>&, int, const int*, const int* const*, \
                                           const int* const*, const T* const*,                  \
                                           const sparsesolv_amd::IcSettings&);
SSA_DECLARE_BATCH(double, double)
SSA_DECLARE_BATCH(double, float)
SSA_DECLARE_BATCH(cd, cd)
SSA_DECLARE_BATCH(cd, cf)
#undef SSA_DECLARE_BATCH

}  // namespace ssa_host
