// Uses the header-only C++ drop-in exactly like reference client code
// (namespace sparsesolv, SolverConfig, solve_iccg, solve_cocr, CGSolver with
// an ICPreconditioner) and prints one JSON line per solve for the GPU test.
#define SPARSESOLV_AMD_AS_SPARSESOLV
#include "sparsesolv_amd/sparsesolv.hpp"

#include <cmath>
#include <cstdio>
#include <vector>

using sparsesolv::index_t;

template <typename T>
static void laplace(int m, std::vector<index_t>& rp, std::vector<index_t>& ci, std::vector<T>& v,
                    T diag)
{
    const int n = m * m * m;
    rp.assign(1, 0);
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
    (void)n;
}

template <typename T>
static double resid(const std::vector<index_t>& rp, const std::vector<index_t>& ci,
                    const std::vector<T>& v, const std::vector<T>& x, const std::vector<T>& b)
{
    double rr = 0, bb = 0;
    for (size_t i = 0; i + 1 < rp.size(); ++i) {
        T s = 0;
        for (index_t k = rp[i]; k < rp[i + 1]; ++k) s += v[k] * x[ci[k]];
        rr += std::norm(b[i] - s);
        bb += std::norm(b[i]);
    }
    return std::sqrt(rr / bb);
}

int main()
{
    const int m = 20;
    {
        std::vector<index_t> rp, ci;
        std::vector<double> v;
        laplace<double>(m, rp, ci, v, 6.0);
        const index_t n = index_t(rp.size()) - 1;
        std::vector<double> b(n), x(n, 0.0);
        for (index_t i = 0; i < n; ++i) b[i] = std::sin(0.1 * i);
        sparsesolv::SparseMatrixView<double> A(n, n, rp.data(), ci.data(), v.data());
        sparsesolv::SolverConfig cfg;
        cfg.use_abmc = true;
        cfg.save_residual_history = true;
        auto r = sparsesolv::solve_iccg(A, b, x, cfg);
        std::printf("{\"case\": \"iccg\", \"converged\": %d, \"iterations\": %d, \"residual\": %.3e, "
                    "\"true_residual\": %.3e, \"history\": %zu}\n",
                    int(r.converged), r.iterations, r.final_residual, resid(rp, ci, v, x, b),
                    r.residual_history.size());
        // explicit preconditioner + solver objects
        sparsesolv::ICPreconditioner<double> pc(1.05);
        pc.set_config(cfg);
        pc.setup(A);
        sparsesolv::CGSolver<double> cg;
        cg.set_config(cfg);
        std::vector<double> x2(n, 0.0);
        auto r2 = cg.solve(A, b.data(), x2.data(), n, &pc);
        double diff = 0;
        for (index_t i = 0; i < n; ++i) diff = std::max(diff, std::abs(x2[i] - x[i]));
        std::printf("{\"case\": \"cg_object\", \"converged\": %d, \"iterations\": %d, \"max_diff\": %.3e}\n",
                    int(r2.converged), r2.iterations, diff);
        bool threw = false;
        try {
            sparsesolv::ICPreconditioner<double> fresh;
            fresh.apply(b.data(), x2.data(), n);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        std::printf("{\"case\": \"apply_before_setup\", \"threw\": %d}\n", int(threw));
    }
    {
        using C = std::complex<double>;
        std::vector<index_t> rp, ci;
        std::vector<C> v;
        laplace<C>(m, rp, ci, v, C(6.0, 0.5));
        const index_t n = index_t(rp.size()) - 1;
        std::vector<C> b(n), x(n, C(0));
        for (index_t i = 0; i < n; ++i) b[i] = C(std::cos(0.05 * i), std::sin(0.02 * i));
        sparsesolv::SparseMatrixView<C> A(n, n, rp.data(), ci.data(), v.data());
        sparsesolv::SolverConfig cfg;
        cfg.use_abmc = true;
        cfg.abmc_num_colors = 8;
        auto r = sparsesolv::solve_cocr(A, b, x, cfg);
        std::printf("{\"case\": \"cocr\", \"converged\": %d, \"iterations\": %d, \"residual\": %.3e, "
                    "\"true_residual\": %.3e}\n",
                    int(r.converged), r.iterations, r.final_residual, resid(rp, ci, v, x, b));

        // COCRSolver::solve(A, ..., &pc) with pc set up on another matrix M
        // (A's pattern, diagonal shifted): the solve is A's, the factor M's
        std::vector<index_t> mrp, mci;
        std::vector<C> mv;
        laplace<C>(m, mrp, mci, mv, C(6.5, 1.5));
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
        double dmax = 0.0;
        for (index_t i = 0; i < n; ++i) dmax = std::max(dmax, std::abs(xm[i] - xa[i]));
        std::printf("{\"case\": \"cocr_other_pre\", \"converged\": %d, \"iterations\": %d, "
                    "\"true_residual\": %.3e, \"own_pre_iterations\": %d, \"max_diff\": %.3e}\n",
                    int(rm.converged), rm.iterations, resid(rp, ci, v, xm, b), ra.iterations, dmax);
    }
    return 0;
}
