// ic_handle.hpp -- what the C ABI (solver.hip) and the handle
// implementations (ic_*.hip, one per precision pair) share: the context,
// the error type, and the handle interface with its factories.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "sparsesolv_amd.h"
#include "sparsesolv_amd/ic_factor.hpp"

struct ssa_context {
    int device = 0;
    hipStream_t stream = nullptr;
    // the device address of ssa_ic_apply_noalloc's scratch array on this
    // context's GPU (resolved once per device by ssa_context_create)
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
};

namespace ssa_impl {

struct SsaError : std::runtime_error {
    int code;
    SsaError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(expr)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw SsaError(SSA_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct IcBase {
    virtual ~IcBase() = default;
    // nrhs vectors of info().n values each, back to back
    virtual void apply(const void* x, void* y, bool on_device, int nrhs) = 0;
    // apply() of one vector (device pointers) with every work vector in the
    // context's scratch array: no device allocation, no stream or event
    virtual void apply_noalloc(const void* x, void* y) = 0;
    virtual void spmv(const void* x, void* y, bool on_device, int nrhs) = 0;
    virtual void set_operator(int n, const int32_t* row_ptr, const int32_t* col_idx,
                              const void* values) = 0;
    virtual void solve(int method, const void* b, void* x, bool on_device, const ssa_config& cfg,
                       ssa_result* results, double* history, int nrhs) = 0;
    virtual int system_count() const = 0;
    virtual void info(ssa_ic_info& out) const = 0;
    virtual void ordering(int32_t* out) const = 0;
    virtual void schedule(int32_t* co, int32_t* bo) const = 0;
    virtual void factor(int32_t* rp, int32_t* ci, void* vals, void* invd) const = 0;
    virtual void profile(int what, int reps, int nrhs, double& avg_us, double& bytes) = 0;
    // kernel class timed inside solves (-1: none) and the totals so far
    virtual void trace(int what) = 0;
    virtual void trace_read(long long& launches, double& us, double& bytes) const = 0;
};

// one handle implementation per (vector, factor) precision pair; values[s]
// points at system s's entries (double or std::complex<double>)
using MakeFn = IcBase* (*)(ssa_context* ctx, int count, const int32_t* n,
                           const int32_t* const* row_ptr, const int32_t* const* col_idx,
                           const void* const* values, const sparsesolv_amd::IcSettings& st);
IcBase* make_f64(ssa_context*, int, const int32_t*, const int32_t* const*, const int32_t* const*,
                 const void* const*, const sparsesolv_amd::IcSettings&);
IcBase* make_f64_mixed(ssa_context*, int, const int32_t*, const int32_t* const*,
                       const int32_t* const*, const void* const*, const sparsesolv_amd::IcSettings&);
IcBase* make_c128(ssa_context*, int, const int32_t*, const int32_t* const*, const int32_t* const*,
                  const void* const*, const sparsesolv_amd::IcSettings&);
IcBase* make_c128_mixed(ssa_context*, int, const int32_t*, const int32_t* const*,
                        const int32_t* const*, const void* const*, const sparsesolv_amd::IcSettings&);

}  // namespace ssa_impl
