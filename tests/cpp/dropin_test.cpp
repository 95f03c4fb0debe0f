// A client of the header-only C++ drop-in written against the reference's
// signatures (namespace sparsesolv; SolverConfig; solve_iccg / solve_cocr /
// solve_sgsmrtr; IterativeSolver and Preconditioner used through base-class
// pointers; ICPreconditioner and SGSPreconditioner; apply_fused_dot; the
// SparseMatrixView accessors and product).  It prints
// one JSON line per case and writes every matrix, right-hand side and
// solution it used as raw little-endian arrays into the directory given as
// argv[1], where the GPU test compares them with the oracle.
#define SPARSESOLV_AMD_AS_SPARSESOLV
#include "sparsesolv_amd/sparsesolv.hpp"

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

using sparsesolv::index_t;
using C = std::complex<double>;

static std::string g_dir = ".";

template <typename T>
static void dump(const std::string& name, const std::vector<T>& v)
{
    const std::string path = g_dir + "/" + name + ".bin";
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    std::fwrite(v.data(), sizeof(T), v.size(), f);
    std::fclose(f);
}

// 7-point operator on an m^3 grid (x fastest), diag on the diagonal, -1 off it
template <typename T>
static void laplace(int m, std::vector<index_t>& rp, std::vector<index_t>& ci, std::vector<T>& v, T diag)
{
    rp.assign(1, 0);
    ci.clear();
    v.clear();
    auto id = [m](int i, int j, int k) { return i + m * (j + m * k); };
    for (int k = 0; k < m; ++k)
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) {
                const int r = id(i, j, k);
                const int nb[7] = {k > 0 ? id(i, j, k - 1) : -1, j > 0 ? id(i, j - 1, k) : -1,
                                   i > 0 ? id(i - 1, j, k) : -1, r,
                                   i < m - 1 ? id(i + 1, j, k) : -1,
                                   j < m - 1 ? id(i, j + 1, k) : -1,
                                   k < m - 1 ? id(i, j, k + 1) : -1};
                for (int q = 0; q < 7; ++q) {
                    if (nb[q] < 0) continue;
                    ci.push_back(nb[q]);
                    v.push_back(nb[q] == r ? diag : T(-1.0));
                }
                rp.push_back(index_t(ci.size()));
            }
}

template <typename T>
static double resid(const sparsesolv::SparseMatrixView<T>& A, const std::vector<T>& x, const std::vector<T>& b)
{
    std::vector<T> ax(b.size());
    A.multiply(x, ax);
    double rr = 0, bb = 0;
    for (size_t i = 0; i < b.size(); ++i) {
        rr += std::norm(b[i] - ax[i]);
        bb += std::norm(b[i]);
    }
    return std::sqrt(rr / bb);
}

// a preconditioner of the client's own (diagonal scaling), with no device
// implementation
class JacobiPre : public sparsesolv::Preconditioner<double> {
public:
    void setup(const sparsesolv::SparseMatrixView<double>& A) override
    {
        d_.resize(A.rows());
        for (index_t i = 0; i < A.rows(); ++i) d_[i] = 1.0 / A.diagonal(i);
        is_setup_ = true;
    }
    void apply(const double* x, double* y, index_t size) const override
    {
        for (index_t i = 0; i < size; ++i) y[i] = d_[i] * x[i];
    }
    std::string name() const override { return "Jacobi"; }

private:
    std::vector<double> d_;
};

int main(int argc, char** argv)
{
    if (argc > 1) g_dir = argv[1];
    const int m = 20;
    {
        std::vector<index_t> rp, ci;
        std::vector<double> v;
        laplace<double>(m, rp, ci, v, 6.0);
        const index_t n = index_t(rp.size()) - 1;
        std::vector<double> b(n), x(n, 0.0);
        for (index_t i = 0; i < n; ++i) b[i] = std::sin(0.1 * i);
        dump("real_rp", rp);
        dump("real_ci", ci);
        dump("real_v", v);
        dump("real_b", b);
        sparsesolv::SparseMatrixView<double> A(n, n, rp.data(), ci.data(), v.data());
        sparsesolv::SolverConfig cfg;
        cfg.use_abmc = true;
        cfg.save_residual_history = true;
        auto r = sparsesolv::solve_iccg(A, b, x, cfg);
        dump("iccg_x", x);
        dump("iccg_hist", r.residual_history);
        std::printf("{\"case\": \"iccg\", \"converged\": %d, \"iterations\": %d, \"residual\": %.17g, "
                    "\"true_residual\": %.3e, \"history\": %zu}\n",
                    int(r.converged), r.iterations, r.final_residual, resid(A, x, b),
                    r.residual_history.size());

        // the preconditioner and the solver through their base classes
        std::unique_ptr<sparsesolv::Preconditioner<double>> pc(new sparsesolv::ICPreconditioner<double>(1.05));
        static_cast<sparsesolv::ICPreconditioner<double>*>(pc.get())->set_config(cfg);
        pc->setup(A);
        std::unique_ptr<sparsesolv::IterativeSolver<double>> cg(new sparsesolv::CGSolver<double>());
        cg->set_config(cfg);
        std::vector<double> x2(n, 0.0);
        auto r2 = cg->solve(A, b.data(), x2.data(), n, pc.get());
        double diff = 0;
        for (index_t i = 0; i < n; ++i) diff = std::max(diff, std::abs(x2[i] - x[i]));
        std::printf("{\"case\": \"cg_base_classes\", \"solver\": \"%s\", \"precond\": \"%s\", "
                    "\"ready\": %d, \"converged\": %d, \"iterations\": %d, \"max_diff\": %.3e}\n",
                    cg->name().c_str(), pc->name().c_str(), int(pc->is_ready()), int(r2.converged),
                    r2.iterations, diff);

        // y = M^-1 b and b . y in one call
        std::vector<double> y(n);
        const double d = pc->apply_fused_dot(b.data(), b.data(), y.data(), n, false);
        dump("fused_y", y);
        std::printf("{\"case\": \"fused_dot\", \"dot\": %.17g}\n", d);

        bool threw = false;
        try {
            sparsesolv::ICPreconditioner<double> fresh;
            fresh.apply(b.data(), x2.data(), n);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        bool threw_solve = false;
        try {
            sparsesolv::ICPreconditioner<double> fresh;
            cg->solve(A, b.data(), x2.data(), n, &fresh);
        } catch (const std::runtime_error&) {
            threw_solve = true;
        }
        std::printf("{\"case\": \"apply_before_setup\", \"threw\": %d, \"solve_threw\": %d}\n",
                    int(threw), int(threw_solve));

        // a preconditioner without a device implementation is refused
        JacobiPre jac;
        jac.setup(A);
        bool refused = false;
        try {
            cg->solve(A, b.data(), x2.data(), n, &jac);
        } catch (const std::invalid_argument&) {
            refused = true;
        }
        std::vector<double> yj(n);
        const double dj = jac.apply_fused_dot(b.data(), b.data(), yj.data(), n, false);
        double dj_want = 0.0;
        for (index_t i = 0; i < n; ++i) dj_want += b[i] * b[i] / A.diagonal(i);
        std::printf("{\"case\": \"host_preconditioner\", \"refused\": %d, \"dot_rel_err\": %.3e}\n",
                    int(refused), std::abs(dj - dj_want) / std::abs(dj_want));

        // the view's accessors and product against the arrays
        double vmax = 0.0;
        int bad = 0;
        std::vector<double> ax(n), ax2(n, 0.0);
        A.multiply(x, ax);
        for (index_t i = 0; i < n; ++i) {
            const auto rr = A.row_range(i);
            if (rr.first != rp[i] || rr.second != rp[i + 1]) ++bad;
            if (A.diagonal(i) != 6.0) ++bad;
            for (index_t k = rp[i]; k < rp[i + 1]; ++k) {
                if (A(i, ci[k]) != v[k]) ++bad;
                ax2[i] += v[k] * x[ci[k]];
            }
            if (i + 3 * m * m < n && A(i, i + 3 * m * m) != 0.0) ++bad;   // not stored
            vmax = std::max(vmax, std::abs(ax[i] - ax2[i]));
        }
        std::printf("{\"case\": \"view\", \"bad\": %d, \"multiply_max_diff\": %.3e, \"nnz\": %d}\n", bad,
                    vmax, int(A.nnz()));

        // the SGS preconditioner through the base classes, and SGS-MRTR
        std::unique_ptr<sparsesolv::Preconditioner<double>> sgs(new sparsesolv::SGSPreconditioner<double>());
        sgs->setup(A);
        std::vector<double> ys(n), xs(n, 0.0);
        sgs->apply(b.data(), ys.data(), n);
        dump("sgs_y", ys);
        auto rs = cg->solve(A, b.data(), xs.data(), n, sgs.get());
        dump("sgs_cg_x", xs);
        std::printf("{\"case\": \"sgs_cg\", \"precond\": \"%s\", \"converged\": %d, \"iterations\": %d, "
                    "\"true_residual\": %.3e}\n",
                    sgs->name().c_str(), int(rs.converged), rs.iterations, resid(A, xs, b));
        std::vector<double> xm(n, 0.0);
        sparsesolv::SolverConfig mc;
        mc.save_residual_history = true;
        auto rm = sparsesolv::solve_sgsmrtr(A, b, xm, mc);
        dump("sgsmrtr_x", xm);
        dump("sgsmrtr_hist", rm.residual_history);
        // SGSMRTRSolver ignores a preconditioner passed to solve()
        sparsesolv::SGSMRTRSolver<double> mrtr;
        mrtr.set_config(mc);
        std::vector<double> xm2(n, 0.0);
        auto rm2 = mrtr.solve(A, b.data(), xm2.data(), n, sgs.get());
        double dm = 0;
        for (index_t i = 0; i < n; ++i) dm = std::max(dm, std::abs(xm2[i] - xm[i]));
        std::printf("{\"case\": \"sgsmrtr\", \"solver\": \"%s\", \"converged\": %d, \"iterations\": %d, "
                    "\"true_residual\": %.3e, \"with_precond_iterations\": %d, \"with_precond_max_diff\": %.3e}\n",
                    mrtr.name().c_str(), int(rm.converged), rm.iterations, resid(A, xm, b), rm2.iterations, dm);
    }
    {
        std::vector<index_t> rp, ci;
        std::vector<C> v;
        laplace<C>(m, rp, ci, v, C(6.0, 0.5));
        const index_t n = index_t(rp.size()) - 1;
        std::vector<C> b(n), x(n, C(0));
        for (index_t i = 0; i < n; ++i) b[i] = C(std::cos(0.05 * i), std::sin(0.02 * i));
        dump("cplx_rp", rp);
        dump("cplx_ci", ci);
        dump("cplx_v", v);
        dump("cplx_b", b);
        sparsesolv::SparseMatrixView<C> A(n, n, rp.data(), ci.data(), v.data());
        sparsesolv::SolverConfig cfg;
        cfg.use_abmc = true;
        cfg.abmc_num_colors = 8;
        auto r = sparsesolv::solve_cocr(A, b, x, cfg);
        dump("cocr_x", x);
        std::printf("{\"case\": \"cocr\", \"converged\": %d, \"iterations\": %d, \"residual\": %.3e, "
                    "\"true_residual\": %.3e}\n",
                    int(r.converged), r.iterations, r.final_residual, resid(A, x, b));

        // COCRSolver::solve(A, ..., &pc) with pc set up on another matrix M
        // (A's pattern, diagonal shifted): the solve is A's, the factor M's
        std::vector<index_t> mrp, mci;
        std::vector<C> mv;
        laplace<C>(m, mrp, mci, mv, C(6.5, 1.5));
        dump("m_v", mv);
        sparsesolv::SparseMatrixView<C> M(n, n, mrp.data(), mci.data(), mv.data());
        sparsesolv::ICPreconditioner<C> pm, pa;
        pm.set_config(cfg);
        pm.setup(M);
        pa.set_config(cfg);
        pa.setup(A);
        sparsesolv::COCRSolver<C> solver;
        solver.set_config(cfg);
        std::vector<C> xm(n, C(0)), xa(n, C(0));
        auto rm = solver.solve(A, b.data(), xm.data(), n, &pm);
        auto ra = solver.solve(A, b.data(), xa.data(), n, &pa);
        dump("cocr_other_pre_x", xm);
        std::printf("{\"case\": \"cocr_other_pre\", \"converged\": %d, \"iterations\": %d, "
                    "\"true_residual\": %.3e, \"own_pre_iterations\": %d}\n",
                    int(rm.converged), rm.iterations, resid(A, xm, b), ra.iterations);

        // A's values changed in place after the preconditioner's setup: the
        // solve multiplies by the live A, the factor stays the set-up one
        for (index_t i = 0; i < n; ++i)
            for (index_t k = rp[i]; k < rp[i + 1]; ++k)
                if (ci[k] == i) v[k] += C(0.3, 0.2);
        dump("cplx_v_new", v);
        std::vector<C> xn(n, C(0));
        auto rn = solver.solve(A, b.data(), xn.data(), n, &pa);
        dump("cocr_changed_x", xn);
        std::printf("{\"case\": \"cocr_changed_values\", \"converged\": %d, \"iterations\": %d, "
                    "\"true_residual\": %.3e}\n",
                    int(rn.converged), rn.iterations, resid(A, xn, b));

        // COCR with the SGS preconditioner, and SGS-MRTR, on the changed A
        sparsesolv::SGSPreconditioner<C> sgs;
        sgs.setup(A);
        std::vector<C> xs(n, C(0)), xg(n, C(0));
        auto rs = solver.solve(A, b.data(), xs.data(), n, &sgs);
        dump("cplx_sgs_cocr_x", xs);
        auto rg = sparsesolv::solve_sgsmrtr(A, b, xg, cfg);
        dump("cplx_sgsmrtr_x", xg);
        std::printf("{\"case\": \"cplx_sgs\", \"cocr_converged\": %d, \"cocr_iterations\": %d, "
                    "\"cocr_true_residual\": %.3e, \"mrtr_converged\": %d, \"mrtr_iterations\": %d}\n",
                    int(rs.converged), rs.iterations, resid(A, xs, b), int(rg.converged), rg.iterations);
    }
    return 0;
}
