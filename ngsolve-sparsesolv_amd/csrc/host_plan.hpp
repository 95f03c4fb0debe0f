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
    double seconds = 0.0;                  // host setup wall time
};

// Throws std::runtime_error (reference messages) on failure.
template <typename T, typename F>
void build_plan(HostPlan<T, F>& plan, int n, const int* row_ptr, const int* col_idx,
                const T* values, const sparsesolv_amd::IcSettings& settings);

using cd = std::complex<double>;
using cf = std::complex<float>;

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

}  // namespace ssa_host
