// sparsesolv.hpp -- header-only C++17 drop-in for the ICCG / COCR / SGS path of
// ksugahar/ngsolve-sparsesolv (include/sparsesolv/sparsesolv.hpp), running on
// one MI355X through the C ABI of libsparsesolv_amd.so.
//
// The names, members and defaults follow the reference:
//   SolverConfig        core/solver_config.hpp:33-159
//   SolverResult        core/types.hpp:29-37
//   SparseMatrixView    core/sparse_matrix_view.hpp:58-173 (non-owning CSR)
//   Preconditioner      core/preconditioner.hpp:22-72
//   ICPreconditioner    preconditioners/ic_preconditioner.hpp:52-1543
//   SGSPreconditioner   preconditioners/sgs_preconditioner.hpp:47-203
//   IterativeSolver     solvers/iterative_solver.hpp:44-322
//   CGSolver, COCRSolver solvers/cg_solver.hpp:59-147, solvers/cocr_solver.hpp:58-187
//   SGSMRTRSolver       solvers/sgs_mrtr_solver.hpp:53-365
//   solve_iccg, solve_cocr, solve_sgsmrtr sparsesolv.hpp:43-105, :153-169, :123-150
// Errors surface as the reference's exception types: std::runtime_error for
// setup failures (missing diagonal) and apply-before-setup,
// std::invalid_argument for bad arguments.
//
// Define SPARSESOLV_AMD_AS_SPARSESOLV before including to also expose the
// classes as namespace `sparsesolv`.
#pragma once

#include <algorithm>
#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../sparsesolv_amd.h"

namespace sparsesolv_amd {

using index_t = std::int32_t;
using complex_t = std::complex<double>;

enum class NormType { RHS, InitialResidual, Custom };
enum class DivergenceCheck { None, StagnationCount };

struct SolverConfig {
    double tolerance = 1e-10;
    int max_iterations = 1000;
    double abs_tolerance = 1e-12;
    NormType norm_type = NormType::RHS;
    double custom_norm = 1.0;
    double shift_parameter = 1.05;
    bool diagonal_scaling = false;
    bool auto_shift = false;
    double shift_increment = 0.05;
    double max_shift_value = 5.0;
    double min_diagonal_threshold = 1e-6;
    double zero_diagonal_replacement = 1e-10;
    int max_shift_trials = 100;
    DivergenceCheck divergence_check = DivergenceCheck::None;
    double divergence_threshold = 1000.0;
    int divergence_count = 100;
    bool save_best_result = true;
    bool save_residual_history = false;
    bool conjugate = false;
    bool use_abmc = false;
    int abmc_block_size = 4;
    int abmc_num_colors = 4;
    bool abmc_reorder_spmv = false;
    bool abmc_use_rcm = false;
    // GPU fields
    int precision = SSA_PRECISION_FP64;
    bool use_graph = true;
    int walk_overlap = -1;
    int rhs_block = 0;
    int device = 0;

    ssa_config to_c() const
    {
        ssa_config c;
        ssa_config_default(&c);
        c.tolerance = tolerance;
        c.max_iterations = max_iterations;
        c.abs_tolerance = abs_tolerance;
        c.norm_type = static_cast<int>(norm_type);
        c.custom_norm = custom_norm;
        c.shift_parameter = shift_parameter;
        c.diagonal_scaling = diagonal_scaling;
        c.auto_shift = auto_shift;
        c.shift_increment = shift_increment;
        c.max_shift_value = max_shift_value;
        c.min_diagonal_threshold = min_diagonal_threshold;
        c.zero_diagonal_replacement = zero_diagonal_replacement;
        c.max_shift_trials = max_shift_trials;
        c.divergence_check = divergence_check == DivergenceCheck::StagnationCount ? 1 : 0;
        c.divergence_threshold = divergence_threshold;
        c.divergence_count = divergence_count;
        c.save_best_result = save_best_result;
        c.save_residual_history = save_residual_history;
        c.conjugate = conjugate;
        c.use_abmc = use_abmc;
        c.abmc_block_size = abmc_block_size;
        c.abmc_num_colors = abmc_num_colors;
        c.abmc_reorder_spmv = abmc_reorder_spmv;
        c.abmc_use_rcm = abmc_use_rcm;
        c.precision = precision;
        c.use_graph = use_graph;
        c.walk_overlap = walk_overlap;
        c.rhs_block = rhs_block;
        return c;
    }
};

struct SolverResult {
    bool converged = false;
    int iterations = 0;
    double final_residual = 0.0;
    std::vector<double> residual_history;
    explicit operator bool() const { return converged; }
};

namespace detail {

inline void check(int rc)
{
    if (rc == SSA_OK) return;
    const std::string msg = ssa_last_error();
    if (rc == SSA_ERR_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

template <typename Scalar>
constexpr int dtype_of()
{
    static_assert(std::is_same_v<Scalar, double> || std::is_same_v<Scalar, complex_t>,
                  "Scalar must be double or std::complex<double>");
    return std::is_same_v<Scalar, double> ? SSA_F64 : SSA_C128;
}

// one context (HIP stream) per device, shared by every object of the process
inline std::shared_ptr<ssa_context> context(int device)
{
    static std::vector<std::weak_ptr<ssa_context>> cache;
    if (device < 0) throw std::invalid_argument("device must be >= 0");
    if (cache.size() <= static_cast<size_t>(device)) cache.resize(device + 1);
    if (auto c = cache[device].lock()) return c;
    ssa_context* raw = nullptr;
    check(ssa_context_create(device, &raw));
    std::shared_ptr<ssa_context> c(raw, [](ssa_context* p) { ssa_context_destroy(p); });
    cache[device] = c;
    return c;
}

}  // namespace detail

// Non-owning CSR view (core/sparse_matrix_view.hpp:58-173): the caller keeps
// the arrays alive.  multiply() is the view's host matrix-vector product, as
// in the reference (client code uses it for residual checks); the solvers
// below run their products on the GPU.
template <typename Scalar = double>
class SparseMatrixView {
public:
    SparseMatrixView() = default;
    SparseMatrixView(index_t rows, index_t cols, const index_t* row_ptr, const index_t* col_idx,
                     const Scalar* values)
        : rows_(rows), cols_(cols), row_ptr_(row_ptr), col_idx_(col_idx), values_(values) {}
    index_t rows() const { return rows_; }
    index_t cols() const { return cols_; }
    index_t nnz() const { return row_ptr_ ? row_ptr_[rows_] : 0; }
    const index_t* row_ptr() const { return row_ptr_; }
    const index_t* col_idx() const { return col_idx_; }
    const Scalar* values() const { return values_; }

    // [start, end) of row i in col_idx / values
    std::pair<index_t, index_t> row_range(index_t i) const
    {
        if (i < 0 || i >= rows_) throw std::out_of_range("SparseMatrixView::row_range: row out of range");
        return {row_ptr_[i], row_ptr_[i + 1]};
    }

    // A(i, j) by binary search of the row's sorted columns; 0 when absent
    Scalar operator()(index_t i, index_t j) const
    {
        const auto r = row_range(i);
        const index_t* it = std::lower_bound(col_idx_ + r.first, col_idx_ + r.second, j);
        if (it != col_idx_ + r.second && *it == j) return values_[it - col_idx_];
        return Scalar(0);
    }

    Scalar diagonal(index_t i) const { return (*this)(i, i); }

    // y = A x (x: cols entries, y: rows entries)
    void multiply(const Scalar* x, Scalar* y) const
    {
        for (index_t i = 0; i < rows_; ++i) {
            Scalar sum = Scalar(0);
            for (index_t k = row_ptr_[i]; k < row_ptr_[i + 1]; ++k) sum += values_[k] * x[col_idx_[k]];
            y[i] = sum;
        }
    }

    void multiply(const std::vector<Scalar>& x, std::vector<Scalar>& y) const
    {
        if (static_cast<index_t>(x.size()) < cols_ || static_cast<index_t>(y.size()) < rows_)
            throw std::invalid_argument("SparseMatrixView::multiply: vector too short");
        multiply(x.data(), y.data());
    }

private:
    index_t rows_ = 0, cols_ = 0;
    const index_t* row_ptr_ = nullptr;
    const index_t* col_idx_ = nullptr;
    const Scalar* values_ = nullptr;
};

namespace detail {

// r . y (conj(r) . y when conjugate), summed in index order
template <typename Scalar>
Scalar dot(const Scalar* r, const Scalar* y, index_t size, bool conjugate)
{
    Scalar s = Scalar(0);
    for (index_t i = 0; i < size; ++i) {
        if constexpr (std::is_same_v<Scalar, double>) {
            (void)conjugate;
            s += r[i] * y[i];
        } else {
            s += (conjugate ? std::conj(r[i]) : r[i]) * y[i];
        }
    }
    return s;
}

// a fingerprint of a matrix's values (whether they changed in place)
template <typename Scalar>
std::uint64_t fingerprint(const Scalar* v, index_t nnz)
{
    const unsigned char* p = reinterpret_cast<const unsigned char*>(v);
    const size_t bytes = static_cast<size_t>(nnz) * sizeof(Scalar);
    std::uint64_t h = 0x9E3779B97F4A7C15ull;
    for (size_t k = 0; k + 8 <= bytes; k += 8) {
        std::uint64_t w;
        std::memcpy(&w, p + k, 8);
        h = (h ^ w) * 0xBF58476D1CE4E5B9ull;
        h ^= h >> 31;
    }
    return h;
}

}  // namespace detail

// Abstract preconditioner (core/preconditioner.hpp:22-72): setup(A), then
// apply(x, y) = M^{-1} x.  device_handle_for() is this path's addition: the
// library handle that runs the preconditioner inside a GPU solve of A, or
// null for a preconditioner with no device implementation (the GPU solvers
// reject those with std::invalid_argument).
template <typename Scalar = double>
class Preconditioner {
public:
    virtual ~Preconditioner() = default;
    virtual void setup(const SparseMatrixView<Scalar>& A) = 0;
    virtual void apply(const Scalar* x, Scalar* y, index_t size) const = 0;

    void apply(const std::vector<Scalar>& x, std::vector<Scalar>& y) const
    {
        if (x.size() != y.size()) throw std::invalid_argument("Preconditioner::apply: size mismatch");
        apply(x.data(), y.data(), static_cast<index_t>(x.size()));
    }

    // y = M^{-1} x and dot(r_for_dot, y) (conj(r_for_dot) . y when
    // conjugate): apply, then the dot (core/preconditioner.hpp:51-65)
    virtual Scalar apply_fused_dot(const Scalar* r_for_dot, const Scalar* x, Scalar* y, index_t size,
                                   bool conjugate) const
    {
        apply(x, y, size);
        return detail::dot(r_for_dot, y, size, conjugate);
    }

    virtual std::string name() const = 0;
    virtual bool is_ready() const { return is_setup_; }

    virtual ssa_ic* device_handle_for(const SparseMatrixView<Scalar>& A) const
    {
        (void)A;
        return nullptr;
    }

protected:
    bool is_setup_ = false;
};

namespace detail {

// A preconditioner whose factor lives on the GPU: setup() builds a library
// handle on A (schedule and factor on the host, uploaded once) from
// handle_config(); apply() runs its sweeps on the GPU (host vectors in, host
// vectors out).
template <typename Scalar>
class DeviceFactor : public Preconditioner<Scalar> {
public:
    void setup(const SparseMatrixView<Scalar>& A) override
    {
        ctx_ = detail::context(device());
        ssa_config c = handle_config();
        ssa_ic* raw = nullptr;
        detail::check(ssa_ic_create(ctx_.get(), detail::dtype_of<Scalar>(), A.rows(), A.row_ptr(),
                                    A.col_idx(), A.values(), &c, &raw));
        ic_.reset(raw, [](ssa_ic* p) { ssa_ic_destroy(p); });
        n_ = A.rows();
        a_ = A.values();
        // kept for solves of another operator with this factor
        const index_t nnz = A.rows() > 0 ? A.row_ptr()[A.rows()] : 0;
        m_rp_.assign(A.row_ptr(), A.row_ptr() + A.rows() + 1);
        m_ci_.assign(A.col_idx(), A.col_idx() + nnz);
        m_v_.assign(A.values(), A.values() + nnz);
        a_print_ = detail::fingerprint(A.values(), nnz);
        op_.reset();
        this->is_setup_ = true;
    }

    // The handle a solve of operator A runs on.  The reference's solvers
    // multiply by the live A at every solve and use the preconditioner as
    // given, whatever matrix it was set up on (solvers/iterative_solver.hpp:
    // 58-80).  A the matrix this preconditioner was set up on: its own
    // handle, whose operator is uploaded again when A's values have changed
    // in place since (the factor stays the set-up one); another matrix of
    // the same order: a second handle with the same factor (set up once on
    // the kept copy of the set-up matrix) whose operator A is uploaded at
    // every call.
    ssa_ic* device_handle_for(const SparseMatrixView<Scalar>& A) const override
    {
        if (!ic_) throw std::runtime_error(this->name() + "Preconditioner::apply called before setup");
        if (A.values() == a_) {
            const std::uint64_t now = detail::fingerprint(A.values(), A.nnz());
            if (now != a_print_) {
                detail::check(ssa_ic_set_operator(ic_.get(), A.rows(), A.row_ptr(), A.col_idx(), A.values()));
                a_print_ = now;
            }
            return ic_.get();
        }
        if (A.rows() != n_) throw std::invalid_argument("solve: the preconditioner's order differs");
        if (!op_) {
            ssa_config c = handle_config();
            ssa_ic* raw = nullptr;
            detail::check(ssa_ic_create(ctx_.get(), detail::dtype_of<Scalar>(), n_, m_rp_.data(),
                                        m_ci_.data(), m_v_.data(), &c, &raw));
            op_.reset(raw, [](ssa_ic* p) { ssa_ic_destroy(p); });
        }
        detail::check(ssa_ic_set_operator(op_.get(), A.rows(), A.row_ptr(), A.col_idx(), A.values()));
        return op_.get();
    }

    void apply(const Scalar* x, Scalar* y, index_t size) const override
    {
        if (!ic_) throw std::runtime_error(this->name() + "Preconditioner::apply called before setup");
        if (size != n_) throw std::invalid_argument(this->name() + "Preconditioner::apply: size mismatch");
        detail::check(ssa_ic_apply(ic_.get(), x, y, 0));
    }
    using Preconditioner<Scalar>::apply;

    // the GPU apply, then r . y on the host (ic_preconditioner.hpp:516-532)
    Scalar apply_fused_dot(const Scalar* r_for_dot, const Scalar* x, Scalar* y, index_t size,
                           bool conjugate) const override
    {
        if (!ic_)
            throw std::runtime_error(this->name() + "Preconditioner::apply_fused_dot called before setup");
        apply(x, y, size);
        return detail::dot(r_for_dot, y, size, conjugate);
    }

    ssa_ic* handle() const { return ic_.get(); }
    const Scalar* matrix_values() const { return a_; }

protected:
    virtual ssa_config handle_config() const = 0;
    virtual int device() const { return 0; }

    std::shared_ptr<ssa_context> ctx_;
    std::shared_ptr<ssa_ic> ic_;
    index_t n_ = 0;

private:
    const Scalar* a_ = nullptr;
    mutable std::uint64_t a_print_ = 0;
    std::vector<index_t> m_rp_, m_ci_;
    std::vector<Scalar> m_v_;
    mutable std::shared_ptr<ssa_ic> op_;
};

// the handle of a solve without a preconditioner (the operator, and for
// SGS-MRTR the solver's own split factor)
template <typename Scalar>
class OperatorHandle final : public DeviceFactor<Scalar> {
public:
    OperatorHandle(const ssa_config& c, int device) : c_(c), device_(device) {}
    std::string name() const override { return "Operator"; }

protected:
    ssa_config handle_config() const override { return c_; }
    int device() const override { return device_; }

private:
    ssa_config c_;
    int device_;
};

}  // namespace detail

// IC(0) preconditioner whose factor lives on the GPU
// (preconditioners/ic_preconditioner.hpp:52-1543): the ABMC schedule and the
// factorisation on the host at setup(), the triangular sweeps on the GPU.
template <typename Scalar = double>
class ICPreconditioner : public detail::DeviceFactor<Scalar> {
public:
    explicit ICPreconditioner(double shift = 1.05) { config_.shift_parameter = shift; }

    void set_config(const SolverConfig& c) { config_ = c; }
    const SolverConfig& config() const { return config_; }
    double shift_parameter() const { return config_.shift_parameter; }
    void set_shift_parameter(double s) { config_.shift_parameter = s; }

    std::string name() const override { return "IC"; }

    double actual_shift() const
    {
        ssa_ic_info info;
        detail::check(ssa_ic_info_get(this->ic_.get(), &info));
        return info.actual_shift;
    }

    std::vector<index_t> abmc_ordering() const
    {
        std::vector<index_t> o(this->n_);
        detail::check(ssa_ic_ordering(this->ic_.get(), o.data()));
        return o;
    }

protected:
    ssa_config handle_config() const override { return config_.to_c(); }
    int device() const override { return config_.device; }

private:
    SolverConfig config_;
};

// Symmetric Gauss-Seidel preconditioner (preconditioners/
// sgs_preconditioner.hpp:47-203): M = (D + L) D^-1 (D + U) from A's own
// triangles and diagonal, applied as two level-schedule sweeps on the GPU.
// No settings (the reference's has none); device 0 unless set.
template <typename Scalar = double>
class SGSPreconditioner : public detail::DeviceFactor<Scalar> {
public:
    SGSPreconditioner() = default;
    void set_device(int d) { device_ = d; }
    std::string name() const override { return "SGS"; }

protected:
    ssa_config handle_config() const override
    {
        ssa_config c;
        ssa_config_default(&c);
        c.preconditioner = SSA_PRECOND_SGS;
        return c;
    }
    int device() const override { return device_; }

private:
    int device_ = 0;
};

// Abstract iterative solver (solvers/iterative_solver.hpp:44-322): solve(A, b,
// x, n, precond) with x the initial guess on entry.  The whole iteration runs
// on the GPU; method() names the library's loop for this solver with (true)
// or without a preconditioner.
template <typename Scalar = double>
class IterativeSolver {
public:
    virtual ~IterativeSolver() = default;

    SolverResult solve(const SparseMatrixView<Scalar>& A, const Scalar* b, Scalar* x, index_t size,
                       const Preconditioner<Scalar>* precond = nullptr)
    {
        if (size != A.rows()) throw std::invalid_argument("solve: size mismatch");
        if (!takes_preconditioner()) precond = nullptr;
        // operator only: a handle on A with the solver's settings
        ssa_config oc = config_.to_c();
        oc.use_abmc = 1;
        oc.preconditioner = own_factor();
        detail::OperatorHandle<Scalar> local(oc, config_.device);
        ssa_ic* h = nullptr;
        if (precond) {
            if (!precond->is_ready())
                throw std::runtime_error(precond->name() + " preconditioner used before setup");
            h = precond->device_handle_for(A);
            if (!h)
                throw std::invalid_argument("IterativeSolver::solve: the " + precond->name() +
                                            " preconditioner has no GPU implementation on this path");
        } else {
            local.setup(A);
            h = local.handle();
        }
        ssa_config c = config_.to_c();
        std::vector<double> hist(static_cast<size_t>(std::max(config_.max_iterations, 0)) + 4);
        ssa_result r;
        detail::check(ssa_solve(h, method(precond != nullptr), b, x, 0, &c, &r, hist.data()));
        SolverResult out;
        out.converged = r.converged != 0;
        out.iterations = r.iterations;
        out.final_residual = r.final_residual;
        if (config_.save_residual_history)
            out.residual_history.assign(hist.begin(), hist.begin() + r.history_len);
        return out;
    }

    SolverResult solve(const SparseMatrixView<Scalar>& A, const std::vector<Scalar>& b,
                       std::vector<Scalar>& x, const Preconditioner<Scalar>* precond = nullptr)
    {
        if (x.size() != b.size()) x.resize(b.size());
        return solve(A, b.data(), x.data(), static_cast<index_t>(b.size()), precond);
    }

    void set_config(const SolverConfig& c) { config_ = c; }
    const SolverConfig& config() const { return config_; }
    SolverConfig& config() { return config_; }
    virtual std::string name() const = 0;

protected:
    virtual int method(bool preconditioned) const = 0;
    // what the handle of a solve without a preconditioner holds besides A
    virtual int own_factor() const { return SSA_PRECOND_IC; }
    virtual bool takes_preconditioner() const { return true; }
    SolverConfig config_;
};

// CG (COCG for complex symmetric, conjugate = false) (solvers/cg_solver.hpp:59-147)
template <typename Scalar = double>
class CGSolver : public IterativeSolver<Scalar> {
public:
    std::string name() const override { return "CG"; }

protected:
    int method(bool pre) const override { return pre ? SSA_METHOD_ICCG : SSA_METHOD_CG; }
};

// COCR (solvers/cocr_solver.hpp:58-187)
template <typename Scalar = double>
class COCRSolver : public IterativeSolver<Scalar> {
public:
    std::string name() const override { return "COCR"; }

protected:
    int method(bool pre) const override { return pre ? SSA_METHOD_ICCOCR : SSA_METHOD_COCR; }
};

// SGS-MRTR (solvers/sgs_mrtr_solver.hpp:53-365): MRTR in the scaled space of
// D A D with the split SGS formula of L = lower(D A D), its own factor (a
// preconditioner passed to solve() is ignored, as in the reference).  FP64,
// one right-hand side.
template <typename Scalar = double>
class SGSMRTRSolver : public IterativeSolver<Scalar> {
public:
    std::string name() const override { return "SGS-MRTR"; }

protected:
    int method(bool) const override { return SSA_METHOD_SGSMRTR; }
    int own_factor() const override { return SSA_PRECOND_SGSMRTR; }
    bool takes_preconditioner() const override { return false; }
};

template <typename Scalar = double>
[[nodiscard]] inline SolverResult solve_iccg(const SparseMatrixView<Scalar>& A, const Scalar* b,
                                             Scalar* x, index_t size,
                                             const SolverConfig& config = SolverConfig())
{
    ICPreconditioner<Scalar> precond(config.shift_parameter);
    precond.set_config(config);
    precond.setup(A);
    CGSolver<Scalar> solver;
    solver.set_config(config);
    return solver.solve(A, b, x, size, &precond);
}

template <typename Scalar = double>
[[nodiscard]] inline SolverResult solve_iccg(const SparseMatrixView<Scalar>& A,
                                             const std::vector<Scalar>& b, std::vector<Scalar>& x,
                                             const SolverConfig& config = SolverConfig())
{
    if (x.size() != b.size()) x.resize(b.size());
    return solve_iccg(A, b.data(), x.data(), static_cast<index_t>(b.size()), config);
}

template <typename Scalar = double>
[[nodiscard]] inline SolverResult solve_cocr(const SparseMatrixView<Scalar>& A, const Scalar* b,
                                             Scalar* x, index_t size,
                                             const SolverConfig& config = SolverConfig())
{
    ICPreconditioner<Scalar> precond(config.shift_parameter);
    precond.set_config(config);
    precond.setup(A);
    COCRSolver<Scalar> solver;
    solver.set_config(config);
    return solver.solve(A, b, x, size, &precond);
}

template <typename Scalar = double>
[[nodiscard]] inline SolverResult solve_cocr(const SparseMatrixView<Scalar>& A,
                                             const std::vector<Scalar>& b, std::vector<Scalar>& x,
                                             const SolverConfig& config = SolverConfig())
{
    if (x.size() != b.size()) x.resize(b.size());
    return solve_cocr(A, b.data(), x.data(), static_cast<index_t>(b.size()), config);
}

// solve_sgsmrtr (sparsesolv.hpp:123-150)
template <typename Scalar = double>
[[nodiscard]] inline SolverResult solve_sgsmrtr(const SparseMatrixView<Scalar>& A, const Scalar* b,
                                                Scalar* x, index_t size,
                                                const SolverConfig& config = SolverConfig())
{
    SGSMRTRSolver<Scalar> solver;
    solver.set_config(config);
    return solver.solve(A, b, x, size);
}

template <typename Scalar = double>
[[nodiscard]] inline SolverResult solve_sgsmrtr(const SparseMatrixView<Scalar>& A,
                                                const std::vector<Scalar>& b, std::vector<Scalar>& x,
                                                const SolverConfig& config = SolverConfig())
{
    if (x.size() != b.size()) x.resize(b.size());
    return solve_sgsmrtr(A, b.data(), x.data(), static_cast<index_t>(b.size()), config);
}

}  // namespace sparsesolv_amd

#ifdef SPARSESOLV_AMD_AS_SPARSESOLV
namespace sparsesolv = sparsesolv_amd;
#endif
