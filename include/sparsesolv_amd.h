/*
 * sparsesolv_amd.h -- C ABI of the MI355X (gfx950) ABMC-IC(0) Krylov path.
 *
 * The library (libsparsesolv_amd.so) replaces, for one GPU, the hot path of
 * ksugahar/ngsolve-sparsesolv: ICPreconditioner on its ABMC branch,
 * CGSolver / COCRSolver driven by it, and the solve_iccg / solve_cocr entry
 * points.  Each function below names the reference interface it stands in
 * for (paths relative to that repository).  Plain pointers and sizes only;
 * every function returns SSA_OK or a negative status, with the message in
 * ssa_last_error().  Exceptions never cross this boundary.
 *
 * Conventions
 *   - indices are int32 (reference core/types.hpp:21), CSR rows sorted by
 *     column, column indices unique within a row, diagonal stored;
 *   - dtype SSA_F64 = double, SSA_C128 = complex<double> as (re, im) pairs;
 *   - vectors passed to apply / solve are in the caller's (original)
 *     ordering; `on_device` = 0 for host memory, 1 for device memory of the
 *     handle's GPU;
 *   - a handle is not re-entrant (one HIP stream per context), like the
 *     reference's preconditioner whose apply writes mutable work buffers
 *     (preconditioners/ic_preconditioner.hpp:556-563).
 */
#ifndef SPARSESOLV_AMD_H
#define SPARSESOLV_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SSA_OK 0
#define SSA_ERR_ARGUMENT (-1)    /* bad argument (reference: invalid_argument / type_error) */
#define SSA_ERR_RUNTIME (-2)     /* reference std::runtime_error cases (missing diagonal, ...) */
#define SSA_ERR_DEVICE (-3)      /* HIP runtime failure */
#define SSA_ERR_UNSUPPORTED (-4) /* option not implemented on this path */

#define SSA_F64 0
#define SSA_C128 1

#define SSA_PRECISION_FP64 0   /* factor, sweeps and vectors in double / complex<double> */
#define SSA_PRECISION_MIXED 1  /* factor and sweeps in float / complex<float>, outer FP64 */

#define SSA_METHOD_ICCG 0        /* CG / COCG preconditioned by IC(0) */
#define SSA_METHOD_ICCOCR 1      /* COCR preconditioned by IC(0) */
#define SSA_METHOD_CG 2          /* CG without preconditioner */
#define SSA_METHOD_COCR 3        /* COCR without preconditioner */
#define SSA_METHOD_NGS_ICCOCR 4  /* COCRSolverNGS loop with the IC preconditioner */
#define SSA_METHOD_SGSMRTR 5     /* SGSMRTRSolver (solve_sgsmrtr); needs an SSA_PRECOND_SGSMRTR handle */

/* What a handle's factor is (ssa_config.preconditioner).  SSA_METHOD_ICCG /
 * SSA_METHOD_ICCOCR run CG / COCR preconditioned by it. */
#define SSA_PRECOND_IC 0         /* IC(0) (ICPreconditioner) */
#define SSA_PRECOND_SGS 1        /* symmetric Gauss-Seidel (SGSPreconditioner) */
#define SSA_PRECOND_SGSMRTR 2    /* the split factor of SGSMRTRSolver: lower(D A D) */

typedef struct ssa_context ssa_context;
typedef struct ssa_ic ssa_ic;

/* sparsesolv::SolverConfig (core/solver_config.hpp:33-159) plus GPU fields.
 * Defaults: ssa_config_default. */
typedef struct ssa_config {
    double tolerance;                 /* 1e-10 */
    int max_iterations;               /* 1000 */
    double abs_tolerance;             /* 1e-12 */
    int norm_type;                    /* 0 RHS, 1 initial residual, 2 custom */
    double custom_norm;               /* 1.0 */
    double shift_parameter;           /* 1.05 */
    int diagonal_scaling;             /* 0; 1: factor of S A S, S = diag(1/sqrt|a_ii|) (per system in a batch) */
    int auto_shift;                   /* 0 */
    double shift_increment;           /* 0.05 */
    double max_shift_value;           /* 5.0 */
    double min_diagonal_threshold;    /* 1e-6 */
    double zero_diagonal_replacement; /* 1e-10 */
    int max_shift_trials;             /* 100 */
    int divergence_check;             /* 0 none, 1 stagnation count */
    double divergence_threshold;      /* 1000.0 */
    int divergence_count;             /* 100 */
    int save_best_result;             /* 1 */
    int save_residual_history;        /* 0 */
    int conjugate;                    /* 0 */
    int use_abmc;                     /* 0 (this path requires 1) */
    int abmc_block_size;              /* 4 */
    int abmc_num_colors;              /* 4 */
    int abmc_reorder_spmv;            /* 0 (GPU Krylov always runs in ABMC order) */
    int abmc_use_rcm;                 /* 0; 1: ABMC of the RCM-reordered matrix */
    /* GPU-only fields */
    int precision;                    /* SSA_PRECISION_FP64 */
    int use_graph;                    /* 1: replay a captured iteration */
    int poll_interval;                /* iterations between host status polls (0: auto) */
    int walk_overlap;                 /* -1 auto: the straggler walks run on a second stream beside
                                         the colours that do not read them, where that pays; 0 off;
                                         1 wherever the schedule allows (read at ssa_ic_create) */
    int preconditioner;               /* SSA_PRECOND_IC; SSA_PRECOND_SGS and SSA_PRECOND_SGSMRTR are
                                         built in the natural row order on the level schedule
                                         (use_abmc, diagonal_scaling and the shift settings do
                                         not apply); SSA_PRECOND_SGSMRTR: FP64, one system per
                                         handle (read at ssa_ic_create) */
    int rhs_block;                    /* 0: 16; 1, 4, 8 or 16: right-hand sides per launch of the
                                         several-right-hand-side calls (a tuning knob: results
                                         are the same bits for every width).  Solves read the
                                         solve's cfg (0 there: the handle's), apply / spmv the
                                         handle's (read at ssa_ic_create) */
} ssa_config;

/* sparsesolv::SolverResult (core/types.hpp:29-37) plus timing. */
typedef struct ssa_result {
    int converged;
    int iterations;
    double final_residual;
    int history_len;
    int status;               /* 1 converged, 2 breakdown, 3 diverged, 4 max iterations, 5 rho breakdown */
    double actual_shift;
    double solve_seconds;     /* wall time of the solve (device inputs resident) */
} ssa_result;

typedef struct ssa_ic_info {
    int n;
    int nnz;
    int nnz_l;                /* entries of the factor L, diagonal included */
    int num_colors;
    int num_blocks;
    int rows_per_lane;        /* regular blocks per lane walk (power of two) */
    int big_blocks;           /* oversized blocks scheduled by levels */
    int max_big_rows;
    int max_levels;
    int shift_trials;
    double actual_shift;
    double setup_seconds;
    size_t device_bytes;
    size_t factor_bytes;      /* bytes of one sweep's factor stream (values+indices) */
    size_t spmv_bytes;        /* bytes of the SpMV operator stream (values+indices) */
    int walk_overlap;         /* -1 not decided yet (decided at the first apply), 0 serial sweeps,
                                 1 straggler walks on a second stream */
} ssa_ic_info;

void ssa_config_default(ssa_config* cfg);
const char* ssa_last_error(void);

/* One context per GPU: device selection + a HIP stream. */
int ssa_context_create(int device, ssa_context** out);
int ssa_context_destroy(ssa_context* ctx);

/* ICPreconditioner<Scalar>::setup (preconditioners/ic_preconditioner.hpp:69-201):
 * ABMC schedule, reordering, IC(0) with auto-shift, transpose -- on the host,
 * then the operator and the factor are uploaded once.  Keeps A (in ABMC
 * order) for solves.  With ctx == NULL the handle is host-only: the setup runs
 * and the schedule / factor queries work, nothing touches a GPU, and apply /
 * spmv / solve return SSA_ERR_ARGUMENT. */
int ssa_ic_create(ssa_context* ctx, int dtype, int n, const int32_t* row_ptr,
                  const int32_t* col_idx, const void* values, const ssa_config* cfg,
                  ssa_ic** out);
int ssa_ic_destroy(ssa_ic* ic);
int ssa_ic_info_get(const ssa_ic* ic, ssa_ic_info* out);

/* ICPreconditioner::abmc_ordering (ic_preconditioner.hpp:260-262): ordering[old] = new. */
int ssa_ic_ordering(const ssa_ic* ic, int32_t* ordering);
/* ABMCSchedule::color_offsets (nc+1) and block_offsets (nb+1) (core/abmc_ordering.hpp:47-53). */
int ssa_ic_schedule(const ssa_ic* ic, int32_t* color_offsets, int32_t* block_offsets);
/* The factor L_ (diagonal last per row) and inv_diag_ (ic_preconditioner.hpp:541-543),
 * as computed on the host before upload, in ABMC order: row_ptr[n+1], col_idx[nnz_l],
 * values[nnz_l], inv_diag[n] (scalars of the handle's dtype). */
int ssa_ic_factor(const ssa_ic* ic, int32_t* row_ptr, int32_t* col_idx, void* values,
                  void* inv_diag);

/* ICPreconditioner::apply (ic_preconditioner.hpp:215-225): y = M^{-1} x, run on the GPU.
 * On an SSA_PRECOND_SGS handle SGSPreconditioner::apply (preconditioners/
 * sgs_preconditioner.hpp:113-131); an SSA_PRECOND_SGSMRTR handle has no apply
 * (SSA_ERR_UNSUPPORTED). */
int ssa_ic_apply(ssa_ic* ic, const void* x, void* y, int on_device);

/* No reference counterpart: ssa_ic_apply for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  x and y are device pointers to one vector each in
 * the caller's ordering; the call makes no device allocation (no hipMalloc,
 * hipFree, stream or event creation) and returns y bit for bit as
 * ssa_ic_apply does.  Its work vectors live in one fixed device array of
 * SSA_NOALLOC_SCRATCH_BYTES per GPU, held from the first ssa_context_create
 * on that GPU until the process ends: a system needs 256-byte-rounded
 * (one stop state per system + n scalars + 2 n preconditioner scalars, plus
 * n more in mixed precision and n more for the scaled level-path apply), and
 * a larger one is rejected with SSA_ERR_ARGUMENT.  All calls on one GPU share
 * the array: calls that overlap on one GPU (several threads, several
 * contexts) race, so the caller serialises them. */
#define SSA_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int ssa_ic_apply_noalloc(ssa_ic* ic, const void* x, void* y);
/* SparseMatrixView::multiply (core/sparse_matrix_view.hpp:139-147): y = A x on the GPU. */
int ssa_spmv(ssa_ic* ic, const void* x, void* y, int on_device);
/* The operator of the handle's solves and SpMV replaced by another matrix of the
 * same order and dtype (any pattern); the preconditioner keeps its factor.  The
 * reference pairs an operator with a preconditioner set up apart:
 * IterativeSolver::solve(A, b, x, n, precond) (solvers/iterative_solver.hpp:58-80)
 * and COCRSolverNGS(mat, pre) (ngsolve/sparsesolv_solvers.hpp:256-396).  One system
 * per handle (SSA_ERR_UNSUPPORTED for a batch); the matrix is copied. */
int ssa_ic_set_operator(ssa_ic* ic, int n, const int32_t* row_ptr, const int32_t* col_idx,
                        const void* values);

/* solve_iccg / solve_cocr (sparsesolv.hpp:43-105, :153-169), CGSolver / COCRSolver
 * without a preconditioner (ngsolve/sparsesolv_solvers.hpp:116-123), the
 * COCRSolverNGS loop (ngsolve/sparsesolv_solvers.hpp:273-369) with this IC
 * preconditioner, CGSolver / COCRSolver with an SGSPreconditioner (an
 * SSA_PRECOND_SGS handle), and solve_sgsmrtr (sparsesolv.hpp:123-137, an
 * SSA_PRECOND_SGSMRTR handle, method SSA_METHOD_SGSMRTR, one right-hand side).
 * x holds the initial guess on entry (zeroed for
 * SSA_METHOD_NGS_ICCOCR) and the solution on exit.  `history` receives
 * result->history_len residuals when cfg->save_residual_history (may be NULL;
 * needs max_iterations+2 entries). */
int ssa_solve(ssa_ic* ic, int method, const void* b, void* x, int on_device,
              const ssa_config* cfg, ssa_result* result, double* history);

/* Several independent systems in one handle.  No reference counterpart:
 * the reference builds one ICPreconditioner per system; this is that setup
 * for `count` systems at once, for callers that solve many systems
 * (parameter studies, frequency sweeps).  System s is n[s], row_ptr[s],
 * col_idx[s], values[s]; all share cfg.  Vectors passed to
 * ssa_ic_apply, ssa_spmv and ssa_batch_solve hold the systems back to back
 * in their own orderings (sum of n[s] entries).  The ordering / schedule /
 * factor queries are for one-system handles only. */
int ssa_batch_create(ssa_context* ctx, int dtype, int count, const int32_t* n,
                     const int32_t* const* row_ptr, const int32_t* const* col_idx,
                     const void* const* values, const ssa_config* cfg, ssa_ic** out);
/* ssa_batch_create with an ABMC block size and colour count of its own for
 * each system (block_size[s], num_colors[s]; either may be NULL for cfg's
 * abmc_block_size / abmc_num_colors).  No reference counterpart: a batch
 * whose parameters are chosen per size class (sparsesolv_amd.autotune). */
int ssa_batch_create_each(ssa_context* ctx, int dtype, int count, const int32_t* n,
                          const int32_t* const* row_ptr, const int32_t* const* col_idx,
                          const void* const* values, const ssa_config* cfg,
                          const int32_t* block_size, const int32_t* num_colors, ssa_ic** out);
/* solve_iccg / solve_cocr (or unpreconditioned CG / COCR) for every system
 * of a handle at once: one sequence of launches covers all systems, each
 * with its own scalars and stop status, and results[s] (count entries) is
 * what ssa_solve gives for system s alone, bit for bit. */
int ssa_batch_solve(ssa_ic* ic, int method, const void* b, void* x, int on_device,
                    const ssa_config* cfg, ssa_result* results);
/* Number of systems of a handle (1 unless made by ssa_batch_create). */
int ssa_ic_systems(const ssa_ic* ic, int* count);

/* Several right-hand sides on one matrix.  No reference counterpart: the
 * reference solves one vector per call; these run nrhs of them on a
 * one-system handle in blocks of up to 16 (1, 4, 8 or 16 per launch, a
 * partial block padded), the SpMV and the colour sweeps loading each matrix
 * and factor entry once for the whole block.  b, x, and the vectors of
 * ssa_ic_apply_multi / ssa_spmv_multi hold nrhs vectors of n values back to
 * back (vector k at offset k*n).  ssa_solve_multi fills results[nrhs]; each
 * right-hand side has its own scalars and stop status, and its result is
 * what ssa_solve gives for it alone, bit for bit.  Methods as ssa_batch_solve
 * (not SSA_METHOD_NGS_ICCOCR). */
int ssa_solve_multi(ssa_ic* ic, int method, int nrhs, const void* b, void* x, int on_device,
                    const ssa_config* cfg, ssa_result* results);
int ssa_ic_apply_multi(ssa_ic* ic, int nrhs, const void* x, void* y, int on_device);
int ssa_spmv_multi(ssa_ic* ic, int nrhs, const void* x, void* y, int on_device);

/* Kernel timing with HIP events on the handle's stream (no reference
 * counterpart: measurement hook for the roofline report).  what:
 *   SSA_PROFILE_SPMV    one SpMV + fused dot over the Krylov vectors;
 *   SSA_PROFILE_FORWARD the forward sweep (all colour phases);
 *   SSA_PROFILE_BACKWARD the backward sweep (all colour phases);
 *   SSA_PROFILE_APPLY   the whole preconditioner apply;
 *   SSA_PROFILE_COLOUR_KERNELS the colour kernel: every colour launched on
 *                       its own (not sharing a small-colour kernel, no
 *                       oversized-block work) runs `reps` times back to back
 *                       between two events; both figures are means over
 *                       those colours, 0 when none qualifies.
 * avg_us = mean wall time of one call (colour kernels: of one launch);
 * bytes = algorithmic bytes one call (launch) must move: every operand once,
 * factor entries with the diagonal at value + int32 index, one int32 row
 * pointer and the input and output entry per row, and (colour kernels) each
 * distinct row of earlier colours the launch reads. */
#define SSA_PROFILE_SPMV 0
#define SSA_PROFILE_FORWARD 1
#define SSA_PROFILE_BACKWARD 2
#define SSA_PROFILE_APPLY 3
#define SSA_PROFILE_COLOUR_KERNELS 4
int ssa_ic_profile(ssa_ic* ic, int what, int reps, double* avg_us, double* bytes);
/* The same timing for a block of nrhs right-hand sides (1..16, padded to the
 * block width); the bytes count every vector operand once per right-hand
 * side and the matrix and factor entries once. */
int ssa_ic_profile_multi(ssa_ic* ic, int what, int nrhs, int reps, double* avg_us, double* bytes);

/* Kernel timing inside solves (no reference counterpart: measurement hook for
 * the roofline report).  ssa_ic_trace(ic, what) selects one kernel class and
 * clears the totals; while a class is selected, ssa_solve / ssa_solve_multi /
 * ssa_batch_solve run their iterations without graph capture and bracket every
 * launch of that class in the iteration loop with two HIP events on the stream
 * it runs on.  ssa_ic_trace_read returns the number of timed launches, the sum
 * of their durations (us) and of their algorithmic bytes (every operand once,
 * as for ssa_ic_profile; the best-x copy is not counted).  SSA_TRACE_OFF turns
 * it off and keeps the totals.  Classes:
 *   SSA_TRACE_SPMV    the SpMV (with its fused dot or convergence check);
 *   SSA_TRACE_COLOUR  the colour kernel k_sweep of colours without oversized
 *                     blocks (both sweeps);
 *   SSA_TRACE_UPDATE  x, r (COCR: and rt; COCR on a block of right-hand sides:
 *                     also the previous iteration's p) update with the |r|^2
 *                     partials;
 *   SSA_TRACE_PQ      the p (COCR: and q; COCR on a block: q only) update with
 *                     the best-x copy;
 *   SSA_TRACE_DOT     the dot partials after the preconditioner. */
#define SSA_TRACE_OFF (-1)
#define SSA_TRACE_SPMV 0
#define SSA_TRACE_COLOUR 1
#define SSA_TRACE_UPDATE 2
#define SSA_TRACE_PQ 3
#define SSA_TRACE_DOT 4
int ssa_ic_trace(ssa_ic* ic, int what);
int ssa_ic_trace_read(const ssa_ic* ic, long long* launches, double* total_us, double* total_bytes);

#ifdef __cplusplus
}
#endif

#endif /* SPARSESOLV_AMD_H */
