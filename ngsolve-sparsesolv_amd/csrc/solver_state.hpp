// solver_state.hpp -- device-resident scalar state of one Krylov solve.
//
// The state mirrors the members of sparsesolv::IterativeSolver that its
// iteration updates (solvers/iterative_solver.hpp:324-350): normaliser,
// best residual, divergence counter, residual history length, plus the
// recurrence scalars of CGSolver / COCRSolver.  It is written only by
// the set-up control kernels and by workgroup 0 of the kernels that fuse a
// control step.
#pragma once

namespace ssa {

enum SolveStatus : int {
    kRunning = 0,
    kConverged = 1,     // check_convergence true after the update
    kBreakdown = 2,     // |pAp| or |mu| < 1e-30 (converged flag as checked)
    kDiverged = 3,      // stagnation counter reached
    kMaxIter = 4,       // loop finished without convergence
    kRhoBreakdown = 5,  // COCR |rho| < 1e-30
};

struct SolverState {
    double rz[2];          // CG: (r, z); COCR: rho
    double rz_prev[2];     // rz of the previous iteration (read by beta)
    double alpha[2];
    double beta[2];
    double denom[2];       // last pAp / mu
    double norm_r;         // last ||r||
    double normalizer;
    double best_residual;
    double final_norm;
    double tol;            // relative tolerance of this pass
    int iter;              // reference loop index of the current iteration
    int status;
    int converged;
    int result_iters;
    int improved;          // last check improved best_residual
    int bad_count;
    int hist_len;
    int maxit;             // iteration budget of this pass
    // mixed-precision COCR: ||r|| and the loop index at the last
    // recomputation of rt = M^{-1} r (or the pass start)
    double replace_norm;
    int replace_iter;
    int replace_pad_;
};

// Per-system settings of one pass (set by the host before the pass): the
// relative tolerance, the iteration budget, and whether the system runs at
// all (an inactive system starts stopped and counts no iteration).
struct PassParams {
    double tolerance;
    int max_iterations;
    int active;
};

struct SolveParams {
    double tolerance;
    double abs_tolerance;
    double custom_norm;
    double divergence_threshold;
    int max_iterations;
    int norm_type;
    int save_best;
    int save_history;
    int divergence_check;
    int divergence_count;
    int conjugate;
    int cocr;
    int ngs;               // COCRSolverNGS criteria (ngsolve/sparsesolv_solvers.hpp:273-369)
    int pad_;
};

}  // namespace ssa
