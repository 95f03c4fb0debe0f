// ic_c128.hip -- the handle for FP64 complex systems.
#include "ic_impl.hpp"

namespace ssa_impl {

IcBase* make_c128(ssa_context* ctx, int count, const int32_t* n, const int32_t* const* rp,
                 const int32_t* const* ci, const void* const* values,
                 const sparsesolv_amd::IcSettings& st)
{
    return make_impl<ssa::c128, ssa::c128, std::complex<double>>(ctx, count, n, rp, ci, values, st);
}

}  // namespace ssa_impl
