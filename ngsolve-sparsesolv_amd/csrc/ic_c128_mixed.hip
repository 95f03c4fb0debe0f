// ic_c128_mixed.hip -- the handle for complex systems with a complex<float> factor.
#include "ic_impl.hpp"

namespace ssa_impl {

IcBase* make_c128_mixed(ssa_context* ctx, int count, const int32_t* n, const int32_t* const* rp,
                       const int32_t* const* ci, const void* const* values,
                       const sparsesolv_amd::IcSettings& st)
{
    return make_impl<ssa::c128, ssa::c64, std::complex<double>>(ctx, count, n, rp, ci, values, st);
}

}  // namespace ssa_impl
