// ic_f64.hip -- the handle for FP64 real systems.
#include "ic_impl.hpp"

namespace ssa_impl {

IcBase* make_f64(ssa_context* ctx, int count, const int32_t* n, const int32_t* const* rp,
                const int32_t* const* ci, const void* const* values,
                const sparsesolv_amd::IcSettings& st)
{
    return make_impl<double, double, double>(ctx, count, n, rp, ci, values, st);
}

}  // namespace ssa_impl
