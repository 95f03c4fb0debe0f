// solver.hip -- C ABI of libsparsesolv_amd.  The handles are built by the
// factories of ic_handle.hpp (one translation unit per precision pair); this
// file checks arguments, maps exceptions to status codes and forwards.
#include <hip/hip_runtime.h>

#include <complex>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "ic_handle.hpp"
#include "sparsesolv_amd.h"

using ssa_impl::IcBase;
using ssa_impl::SsaError;

namespace {

thread_local std::string g_last_error;

template <typename Fn>
int guarded(Fn&& fn)
{
    try {
        fn();
        return SSA_OK;
    } catch (const SsaError& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "out of host memory";
        return SSA_ERR_RUNTIME;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return SSA_ERR_RUNTIME;
    }
}

// Work space of ssa_ic_apply_noalloc: a file-scope device array, one copy
// per GPU for the life of the process.  Host code never uses the array's
// name as an address: ssa_context_create resolves the device address once
// per GPU (with that GPU current) and every use goes through it.
static __device__ __align__(256) unsigned char scratch[SSA_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// the scratch array's address on `device` (the current device)
void* scratch_address(int device)
{
    if (device < 0 || device >= kMaxDevices)
        throw SsaError(SSA_ERR_ARGUMENT, "ssa_context_create: device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw SsaError(SSA_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = ptr;
    }
    return g_scratch_ptr[device];
}

}  // namespace

struct ssa_ic {
    std::unique_ptr<IcBase> impl;
    int dtype = 0;
};

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

void ssa_config_default(ssa_config* c)
{
    std::memset(c, 0, sizeof(*c));
    c->tolerance = 1e-10;
    c->max_iterations = 1000;
    c->abs_tolerance = 1e-12;
    c->norm_type = 0;
    c->custom_norm = 1.0;
    c->shift_parameter = 1.05;
    c->diagonal_scaling = 0;
    c->auto_shift = 0;
    c->shift_increment = 0.05;
    c->max_shift_value = 5.0;
    c->min_diagonal_threshold = 1e-6;
    c->zero_diagonal_replacement = 1e-10;
    c->max_shift_trials = 100;
    c->divergence_check = 0;
    c->divergence_threshold = 1000.0;
    c->divergence_count = 100;
    c->save_best_result = 1;
    c->save_residual_history = 0;
    c->conjugate = 0;
    c->use_abmc = 0;
    c->abmc_block_size = 4;
    c->abmc_num_colors = 4;
    c->abmc_reorder_spmv = 0;
    c->abmc_use_rcm = 0;
    c->precision = SSA_PRECISION_FP64;
    c->use_graph = 1;
    c->poll_interval = 0;
    c->walk_overlap = -1;
    c->preconditioner = SSA_PRECOND_IC;
    c->rhs_block = 0;
}

const char* ssa_last_error(void) { return g_last_error.c_str(); }

int ssa_context_create(int device, ssa_context** out)
{
    return guarded([&] {
        if (!out) throw SsaError(SSA_ERR_ARGUMENT, "ssa_context_create: null output");
        int count = 0;
        HIP_CHECK(hipGetDeviceCount(&count));
        if (device < 0 || device >= count)
            throw SsaError(SSA_ERR_ARGUMENT, "ssa_context_create: no such device " + std::to_string(device));
        HIP_CHECK(hipSetDevice(device));
        void* scr = scratch_address(device);
        auto* c = new ssa_context();
        c->device = device;
        c->scratch = scr;
        c->scratch_bytes = SSA_NOALLOC_SCRATCH_BYTES;
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            throw SsaError(SSA_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
        }
        *out = c;
    });
}

int ssa_context_destroy(ssa_context* c)
{
    return guarded([&] {
        if (!c) return;
        (void)hipSetDevice(c->device);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
    });
}

static void check_matrix(const char* who, int n, const int32_t* row_ptr, const int32_t* col_idx,
                         const void* values)
{
    if (n < 0 || (n > 0 && (!row_ptr || !col_idx || !values)))
        throw SsaError(SSA_ERR_ARGUMENT, std::string(who) + ": bad matrix");
    if (n > 0 && row_ptr[0] != 0) throw SsaError(SSA_ERR_ARGUMENT, std::string(who) + ": row_ptr[0] != 0");
    for (int i = 0; i < n; ++i)
        if (row_ptr[i + 1] < row_ptr[i])
            throw SsaError(SSA_ERR_ARGUMENT, std::string(who) + ": row_ptr not monotone");
    const int nnz = n > 0 ? row_ptr[n] : 0;
    for (int k = 0; k < nnz; ++k)
        if (col_idx[k] < 0 || col_idx[k] >= n)
            throw SsaError(SSA_ERR_ARGUMENT, std::string(who) + ": column index out of range");
}

static void check_rhs_block(const char* who, int b)
{
    if (b != 0 && b != 1 && b != 4 && b != 8 && b != 16)
        throw SsaError(SSA_ERR_ARGUMENT, std::string(who) + ": rhs_block must be 0, 1, 4, 8 or 16");
}

static ssa_ic* create_handle(const char* who, ssa_context* ctx, int dtype, int count,
                             const int32_t* n, const int32_t* const* row_ptr,
                             const int32_t* const* col_idx, const void* const* values,
                             const ssa_config* cfg, const int32_t* block_size = nullptr,
                             const int32_t* num_colors = nullptr)
{
    if (!cfg) throw SsaError(SSA_ERR_ARGUMENT, std::string(who) + ": null argument");
    if (dtype != SSA_F64 && dtype != SSA_C128)
        throw SsaError(SSA_ERR_ARGUMENT, std::string(who) + ": dtype must be SSA_F64 or SSA_C128");
    if (cfg->precision != SSA_PRECISION_FP64 && cfg->precision != SSA_PRECISION_MIXED)
        throw SsaError(SSA_ERR_ARGUMENT, std::string(who) + ": unknown precision");
    const int kind = cfg->preconditioner;
    if (kind != SSA_PRECOND_IC && kind != SSA_PRECOND_SGS && kind != SSA_PRECOND_SGSMRTR)
        throw SsaError(SSA_ERR_ARGUMENT, std::string(who) + ": unknown preconditioner");
    if (kind == SSA_PRECOND_SGSMRTR && cfg->precision != SSA_PRECISION_FP64)
        throw SsaError(SSA_ERR_UNSUPPORTED, std::string(who) + ": SGS-MRTR runs in FP64");
    if (kind == SSA_PRECOND_SGSMRTR && count > 1)
        throw SsaError(SSA_ERR_UNSUPPORTED, std::string(who) + ": SGS-MRTR runs one system per handle");
    for (int q = 0; q < count; ++q) check_matrix(who, n[q], row_ptr[q], col_idx[q], values[q]);
    if (ctx) HIP_CHECK(hipSetDevice(ctx->device));
    sparsesolv_amd::IcSettings st;
    st.shift = cfg->shift_parameter;
    st.diagonal_scaling = cfg->diagonal_scaling != 0;
    st.auto_shift = cfg->auto_shift != 0;
    st.shift_increment = cfg->shift_increment;
    st.max_shift_value = cfg->max_shift_value;
    st.min_diagonal_threshold = cfg->min_diagonal_threshold;
    st.zero_diagonal_replacement = cfg->zero_diagonal_replacement;
    st.max_shift_trials = cfg->max_shift_trials;
    st.use_abmc = cfg->use_abmc != 0;
    st.abmc_block_size = cfg->abmc_block_size;
    st.abmc_num_colors = cfg->abmc_num_colors;
    st.abmc_use_rcm = cfg->abmc_use_rcm != 0;
    st.walk_overlap = cfg->walk_overlap < 0 ? -1 : cfg->walk_overlap > 0 ? 1 : 0;
    check_rhs_block(who, cfg->rhs_block);
    st.rhs_block = cfg->rhs_block;
    for (int q = 0; q < count; ++q) {
        if ((block_size && block_size[q] < 1) || (num_colors && num_colors[q] < 1))
            throw SsaError(SSA_ERR_ARGUMENT, std::string(who) + ": block sizes and colour counts must be >= 1");
    }
    st.batch_block_size = block_size;
    st.batch_num_colors = num_colors;
    if (kind != SSA_PRECOND_IC) {
        // the split factors: natural order on the level schedule, no
        // scaling or shift (SGSPreconditioner::setup, SGSMRTRSolver)
        st.factor_kind = kind == SSA_PRECOND_SGS ? sparsesolv_amd::kFactorSgs : sparsesolv_amd::kFactorSgsMrtr;
        st.use_abmc = false;
        st.abmc_use_rcm = false;
        st.diagonal_scaling = false;
        st.auto_shift = false;
    }
    auto h = std::make_unique<ssa_ic>();
    h->dtype = dtype;
    // mixed: the factor (computed in FP64) is stored and applied in
    // float / complex<float>; SpMV, vectors and dots stay FP64
    const bool mixed = cfg->precision == SSA_PRECISION_MIXED;
    const ssa_impl::MakeFn make = dtype == SSA_F64 ? (mixed ? ssa_impl::make_f64_mixed : ssa_impl::make_f64)
                                                   : (mixed ? ssa_impl::make_c128_mixed : ssa_impl::make_c128);
    h->impl.reset(make(ctx, count, n, row_ptr, col_idx, values, st));
    return h.release();
}

int ssa_ic_create(ssa_context* ctx, int dtype, int n, const int32_t* row_ptr,
                  const int32_t* col_idx, const void* values, const ssa_config* cfg, ssa_ic** out)
{
    return guarded([&] {
        if (!out) throw SsaError(SSA_ERR_ARGUMENT, "ssa_ic_create: null argument");
        const int32_t nn = n;
        *out = create_handle("ssa_ic_create", ctx, dtype, 1, &nn, &row_ptr, &col_idx, &values, cfg);
    });
}

int ssa_batch_create(ssa_context* ctx, int dtype, int count, const int32_t* n,
                     const int32_t* const* row_ptr, const int32_t* const* col_idx,
                     const void* const* values, const ssa_config* cfg, ssa_ic** out)
{
    return guarded([&] {
        if (!out || !n || !row_ptr || !col_idx || !values || count < 1)
            throw SsaError(SSA_ERR_ARGUMENT, "ssa_batch_create: bad argument");
        *out = create_handle("ssa_batch_create", ctx, dtype, count, n, row_ptr, col_idx, values, cfg);
    });
}

int ssa_batch_create_each(ssa_context* ctx, int dtype, int count, const int32_t* n,
                          const int32_t* const* row_ptr, const int32_t* const* col_idx,
                          const void* const* values, const ssa_config* cfg,
                          const int32_t* block_size, const int32_t* num_colors, ssa_ic** out)
{
    return guarded([&] {
        if (!out || !n || !row_ptr || !col_idx || !values || count < 1)
            throw SsaError(SSA_ERR_ARGUMENT, "ssa_batch_create_each: bad argument");
        *out = create_handle("ssa_batch_create_each", ctx, dtype, count, n, row_ptr, col_idx, values,
                             cfg, block_size, num_colors);
    });
}

int ssa_ic_systems(const ssa_ic* ic, int* count)
{
    return guarded([&] {
        if (!ic || !count) throw SsaError(SSA_ERR_ARGUMENT, "ssa_ic_systems: null argument");
        *count = ic->impl->system_count();
    });
}

int ssa_ic_destroy(ssa_ic* ic)
{
    return guarded([&] { delete ic; });
}

int ssa_ic_info_get(const ssa_ic* ic, ssa_ic_info* out)
{
    return guarded([&] {
        if (!ic || !out) throw SsaError(SSA_ERR_ARGUMENT, "ssa_ic_info_get: null argument");
        std::memset(out, 0, sizeof(*out));
        ic->impl->info(*out);
    });
}

int ssa_ic_ordering(const ssa_ic* ic, int32_t* ordering)
{
    return guarded([&] {
        if (!ic || !ordering) throw SsaError(SSA_ERR_ARGUMENT, "ssa_ic_ordering: null argument");
        ic->impl->ordering(ordering);
    });
}

int ssa_ic_schedule(const ssa_ic* ic, int32_t* color_offsets, int32_t* block_offsets)
{
    return guarded([&] {
        if (!ic) throw SsaError(SSA_ERR_ARGUMENT, "ssa_ic_schedule: null argument");
        ic->impl->schedule(color_offsets, block_offsets);
    });
}

int ssa_ic_factor(const ssa_ic* ic, int32_t* row_ptr, int32_t* col_idx, void* values, void* inv_diag)
{
    return guarded([&] {
        if (!ic) throw SsaError(SSA_ERR_ARGUMENT, "ssa_ic_factor: null argument");
        ic->impl->factor(row_ptr, col_idx, values, inv_diag);
    });
}

int ssa_ic_apply(ssa_ic* ic, const void* x, void* y, int on_device)
{
    return guarded([&] {
        if (!ic || !x || !y) throw SsaError(SSA_ERR_ARGUMENT, "ssa_ic_apply: null argument");
        ic->impl->apply(x, y, on_device != 0, 1);
    });
}

int ssa_ic_apply_noalloc(ssa_ic* ic, const void* x, void* y)
{
    return guarded([&] {
        if (!ic || !x || !y) throw SsaError(SSA_ERR_ARGUMENT, "ssa_ic_apply_noalloc: null argument");
        ic->impl->apply_noalloc(x, y);
    });
}

int ssa_ic_apply_multi(ssa_ic* ic, int nrhs, const void* x, void* y, int on_device)
{
    return guarded([&] {
        if (!ic || !x || !y) throw SsaError(SSA_ERR_ARGUMENT, "ssa_ic_apply_multi: null argument");
        ic->impl->apply(x, y, on_device != 0, nrhs);
    });
}

int ssa_spmv(ssa_ic* ic, const void* x, void* y, int on_device)
{
    return guarded([&] {
        if (!ic || !x || !y) throw SsaError(SSA_ERR_ARGUMENT, "ssa_spmv: null argument");
        ic->impl->spmv(x, y, on_device != 0, 1);
    });
}

int ssa_ic_set_operator(ssa_ic* ic, int n, const int32_t* row_ptr, const int32_t* col_idx,
                        const void* values)
{
    return guarded([&] {
        if (!ic) throw SsaError(SSA_ERR_ARGUMENT, "ssa_ic_set_operator: null argument");
        check_matrix("ssa_ic_set_operator", n, row_ptr, col_idx, values);
        ic->impl->set_operator(n, row_ptr, col_idx, values);
    });
}

int ssa_spmv_multi(ssa_ic* ic, int nrhs, const void* x, void* y, int on_device)
{
    return guarded([&] {
        if (!ic || !x || !y) throw SsaError(SSA_ERR_ARGUMENT, "ssa_spmv_multi: null argument");
        ic->impl->spmv(x, y, on_device != 0, nrhs);
    });
}

int ssa_solve(ssa_ic* ic, int method, const void* b, void* x, int on_device, const ssa_config* cfg,
              ssa_result* result, double* history)
{
    return guarded([&] {
        if (!ic || !b || !x || !cfg || !result) throw SsaError(SSA_ERR_ARGUMENT, "ssa_solve: null argument");
        if (method < SSA_METHOD_ICCG || method > SSA_METHOD_SGSMRTR)
            throw SsaError(SSA_ERR_ARGUMENT, "ssa_solve: unknown method");
        if (ic->impl->system_count() != 1)
            throw SsaError(SSA_ERR_ARGUMENT, "ssa_solve: a batched handle is solved with ssa_batch_solve");
        std::memset(result, 0, sizeof(*result));
        ic->impl->solve(method, b, x, on_device != 0, *cfg, result, history, 1);
    });
}

int ssa_solve_multi(ssa_ic* ic, int method, int nrhs, const void* b, void* x, int on_device,
                    const ssa_config* cfg, ssa_result* results)
{
    return guarded([&] {
        if (!ic || !b || !x || !cfg || !results)
            throw SsaError(SSA_ERR_ARGUMENT, "ssa_solve_multi: null argument");
        if (method < SSA_METHOD_ICCG || method > SSA_METHOD_COCR)
            throw SsaError(SSA_ERR_ARGUMENT, "ssa_solve_multi: unknown method");
        if (nrhs < 1) throw SsaError(SSA_ERR_ARGUMENT, "ssa_solve_multi: nrhs must be at least 1");
        if (ic->impl->system_count() != 1)
            throw SsaError(SSA_ERR_ARGUMENT, "ssa_solve_multi: a batched handle is solved with ssa_batch_solve");
        check_rhs_block("ssa_solve_multi", cfg->rhs_block);
        std::memset(results, 0, sizeof(ssa_result) * size_t(nrhs));
        ic->impl->solve(method, b, x, on_device != 0, *cfg, results, nullptr, nrhs);
    });
}

int ssa_batch_solve(ssa_ic* ic, int method, const void* b, void* x, int on_device,
                    const ssa_config* cfg, ssa_result* results)
{
    return guarded([&] {
        if (!ic || !b || !x || !cfg || !results) throw SsaError(SSA_ERR_ARGUMENT, "ssa_batch_solve: null argument");
        if (method < SSA_METHOD_ICCG || method > SSA_METHOD_COCR)
            throw SsaError(SSA_ERR_ARGUMENT, "ssa_batch_solve: unknown method");
        const int count = ic->impl->system_count();
        std::memset(results, 0, sizeof(ssa_result) * size_t(count));
        ic->impl->solve(method, b, x, on_device != 0, *cfg, results, nullptr, 1);
    });
}

int ssa_ic_profile(ssa_ic* ic, int what, int reps, double* avg_us, double* bytes)
{
    return guarded([&] {
        if (!ic || !avg_us || !bytes) throw SsaError(SSA_ERR_ARGUMENT, "ssa_ic_profile: null argument");
        if (what < SSA_PROFILE_SPMV || what > SSA_PROFILE_COLOUR_KERNELS)
            throw SsaError(SSA_ERR_ARGUMENT, "ssa_ic_profile: unknown kernel");
        ic->impl->profile(what, reps, 1, *avg_us, *bytes);
    });
}

int ssa_ic_profile_multi(ssa_ic* ic, int what, int nrhs, int reps, double* avg_us, double* bytes)
{
    return guarded([&] {
        if (!ic || !avg_us || !bytes) throw SsaError(SSA_ERR_ARGUMENT, "ssa_ic_profile_multi: null argument");
        if (what < SSA_PROFILE_SPMV || what > SSA_PROFILE_COLOUR_KERNELS)
            throw SsaError(SSA_ERR_ARGUMENT, "ssa_ic_profile_multi: unknown kernel");
        ic->impl->profile(what, reps, nrhs, *avg_us, *bytes);
    });
}

int ssa_ic_trace(ssa_ic* ic, int what)
{
    return guarded([&] {
        if (!ic) throw SsaError(SSA_ERR_ARGUMENT, "ssa_ic_trace: null argument");
        ic->impl->trace(what);
    });
}

int ssa_ic_trace_read(const ssa_ic* ic, long long* launches, double* total_us, double* total_bytes)
{
    return guarded([&] {
        if (!ic || !launches || !total_us || !total_bytes)
            throw SsaError(SSA_ERR_ARGUMENT, "ssa_ic_trace_read: null argument");
        ic->impl->trace_read(*launches, *total_us, *total_bytes);
    });
}

}  // extern "C"
