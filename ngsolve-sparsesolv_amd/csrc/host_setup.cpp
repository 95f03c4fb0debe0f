// host_setup.cpp -- host half of ssa_ic_create: ABMC schedule, IC(0)
// factor and HBM layouts (compiled by g++ with -ffp-contract=off).
#include <chrono>

#include "host_plan.hpp"

namespace ssa_host {

template <typename F, typename T>
static F narrow(const T& v) { return static_cast<F>(v); }
template <>
cf narrow<cf, cd>(const cd& v) { return cf(float(v.real()), float(v.imag())); }

template <typename T, typename F>
void build_plan(HostPlan<T, F>& plan, int n, const int* row_ptr, const int* col_idx,
                const T* values, const sparsesolv_amd::IcSettings& settings)
{
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    plan.fac.build(n, row_ptr, col_idx, values, settings, /*keep_permuted_matrix=*/true);
    plan.spmv.build(plan.fac.A_perm);
    // the sweep layouts are built from the factor converted to F
    sparsesolv_amd::CsrHost<F> L, U;
    L.n = U.n = n;
    L.rp = plan.fac.L.rp; L.ci = plan.fac.L.ci;
    U.rp = plan.fac.U.rp; U.ci = plan.fac.U.ci;
    L.v.resize(plan.fac.L.v.size());
    U.v.resize(plan.fac.U.v.size());
    for (size_t k = 0; k < L.v.size(); ++k) L.v[k] = narrow<F>(plan.fac.L.v[k]);
    for (size_t k = 0; k < U.v.size(); ++k) U.v[k] = narrow<F>(plan.fac.U.v[k]);
    plan.invd.resize(n);
    for (int i = 0; i < n; ++i) plan.invd[i] = narrow<F>(plan.fac.inv_diag[i]);
    plan.fwd.build(plan.fac.sched, L, /*forward=*/true, settings.abmc_block_size, plan.invd);
    plan.bwd.build(plan.fac.sched, U, /*forward=*/false, settings.abmc_block_size, plan.invd);
    plan.seconds = std::chrono::duration<double>(clk::now() - t0).count();
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

}  // namespace ssa_host
