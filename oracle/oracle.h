/*
 * oracle.h -- CPU restatement of the reference ABMC-IC(0) Krylov path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the GPU
 * product is compared against; it is never linked into, loaded by, or
 * called from the product library.  Only tests/, the smoke() entry point
 * and bench.py's cpu_baseline leg may load it.
 *
 * Every function restates the serial (single-thread) behaviour of the
 * reference header library include/sparsesolv (ksugahar/ngsolve-sparsesolv
 * v2.7.0); each definition in oracle_impl.inc cites the file:line it follows.
 * The same entry points are exported by oracle/_ref/libssref.so, which is the
 * reference headers themselves compiled behind oracle/ref_harness.cpp, so a
 * test can run both implementations through one ctypes wrapper.
 *
 * Scalars: dtype 0 = double, dtype 1 = complex<double> stored as (re, im)
 * pairs.  Indices are int32 (reference core/types.hpp:21).
 */
#ifndef SSA_ORACLE_H
#define SSA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Mirror of sparsesolv::SolverConfig (core/solver_config.hpp:33-159).
 * Field order is part of the ABI of this file (ctypes mirrors it). */
typedef struct orc_config {
    double tolerance;                 /* :39  1e-10 */
    int    max_iterations;            /* :42  1000 */
    double abs_tolerance;             /* :45  1e-12 */
    int    norm_type;                 /* :52  0=RHS 1=InitialResidual 2=Custom */
    double custom_norm;               /* :55  1.0 */
    double shift_parameter;           /* :64  1.05 */
    int    diagonal_scaling;          /* :67  0 */
    int    auto_shift;                /* :75  0 */
    double shift_increment;           /* :79  0.05 */
    double max_shift_value;           /* :82  5.0 */
    double min_diagonal_threshold;    /* :85  1e-6 */
    double zero_diagonal_replacement; /* :88  1e-10 */
    int    max_shift_trials;          /* :91  100 */
    int    divergence_check;          /* :98  0=None 1=StagnationCount */
    double divergence_threshold;      /* :101 1000 */
    int    divergence_count;          /* :104 100 */
    int    save_best_result;          /* :111 1 */
    int    save_residual_history;     /* :114 0 */
    int    conjugate;                 /* :123 0 */
    int    use_abmc;                  /* :133 0 */
    int    abmc_block_size;           /* :138 4 */
    int    abmc_num_colors;           /* :144 4 */
    int    abmc_reorder_spmv;         /* :151 0 */
    int    abmc_use_rcm;              /* :157 0 (not supported: returns error) */
} orc_config;

/* Mirror of sparsesolv::SolverResult (core/types.hpp:29-37). */
typedef struct orc_result {
    int    converged;
    int    iterations;
    double final_residual;
    int    history_len;      /* entries written to the caller's history buffer */
    double actual_shift;     /* shift the IC factorisation ended with */
} orc_result;

void orc_config_default(orc_config* c);

/* ---- ABMC schedule (core/abmc_ordering.hpp:83-121) ------------------- */
/* Returns an opaque handle, NULL on bad input. */
void* orc_abmc_build(int n, const int* row_ptr, const int* col_idx,
                     int block_size, int target_colors);
/* which: 0=num_colors 1=num_blocks */
int   orc_abmc_count(const void* h, int which);
/* which: 0=color_offsets[nc+1] 1=color_blocks[nb] 2=block_offsets[nb+1]
 *        3=ordering[n] 4=reverse_ordering[n] */
void  orc_abmc_copy(const void* h, int which, int* out);
void  orc_abmc_free(void* h);

/* ---- IC(0) preconditioner (preconditioners/ic_preconditioner.hpp) ---- */
/* setup (ic_preconditioner.hpp:69-201).  Returns NULL and fills err on
 * failure (missing diagonal). */
void* orc_ic_setup(int dtype, int n, const int* row_ptr, const int* col_idx,
                   const void* values, const orc_config* cfg,
                   char* err, int errlen);
/* which: 0=n 1=nnz_L 2=num_colors(0 if not ABMC) 3=num_blocks */
int    orc_ic_count(const void* h, int which);
double orc_ic_actual_shift(const void* h);
/* which: 0=L.row_ptr 1=L.col_idx 2=Lt.row_ptr 3=Lt.col_idx 4=ordering
 *        5=color_offsets 6=block_offsets */
void   orc_ic_copy_int(const void* h, int which, int* out);
/* which: 0=L.values 1=Lt.values 2=inv_diag 3=scaling (double) */
void   orc_ic_copy_val(const void* h, int which, void* out);
/* y = M^{-1} x in the caller's (original) ordering (apply, :215-225). */
void   orc_ic_apply(const void* h, const void* x, void* y);
/* y = M^{-1} x with x, y already in ABMC order
 * (apply_in_reordered_space, :299-318). */
void   orc_ic_apply_reordered(const void* h, const void* x, void* y);
void   orc_ic_free(void* h);

/* ---- Solvers (sparsesolv.hpp:43-105, :153-169) ----------------------- */
/* x is the initial guess on entry and the solution on exit.  history must
 * hold max_iterations+2 doubles (or be NULL).  Returns 0, or -1 with err
 * filled (the reference throws std::runtime_error in those cases). */
int orc_solve_iccg(int dtype, int n, const int* row_ptr, const int* col_idx,
                   const void* values, const void* b, void* x,
                   const orc_config* cfg, orc_result* res, double* history,
                   char* err, int errlen);
int orc_solve_cocr(int dtype, int n, const int* row_ptr, const int* col_idx,
                   const void* values, const void* b, void* x,
                   const orc_config* cfg, orc_result* res, double* history,
                   char* err, int errlen);
/* Unpreconditioned CG / COCR (SparseSolvSolver method "CG"/"COCR",
 * ngsolve/sparsesolv_solvers.hpp:116-123). */
/* COCR on the operator (row_ptr, col_idx, values) with the IC
   preconditioner set up on another matrix m_* of the same order. */
int orc_solve_cocr_pre(int dtype, int n, const int* row_ptr, const int* col_idx,
                       const void* values, const int* m_row_ptr, const int* m_col_idx,
                       const void* m_values, const void* b, void* x,
                       const orc_config* cfg, orc_result* res, double* history,
                       char* err, int errlen);
int orc_solve_plain(int dtype, int method_cocr, int n, const int* row_ptr,
                    const int* col_idx, const void* values, const void* b,
                    void* x, const orc_config* cfg, orc_result* res,
                    double* history, char* err, int errlen);

/* ---- SGS preconditioner and SGS-MRTR --------------------------------- */
/* y = M^{-1} x of the SGSPreconditioner set up on the matrix
 * (preconditioners/sgs_preconditioner.hpp:71-131). */
void orc_sgs_apply(int dtype, int n, const int* row_ptr, const int* col_idx,
                   const void* values, const void* x, void* y);
/* CGSolver (method_cocr 0) / COCRSolver (1) with an SGSPreconditioner set up
 * on the operator (iterative_solver.hpp:58-80). */
int orc_solve_sgs(int dtype, int method_cocr, int n, const int* row_ptr, const int* col_idx,
                  const void* values, const void* b, void* x, const orc_config* cfg,
                  orc_result* res, double* history, char* err, int errlen);
/* solve_sgsmrtr (sparsesolv.hpp:124-137): SGSMRTRSolver. */
int orc_solve_sgsmrtr(int dtype, int n, const int* row_ptr, const int* col_idx,
                      const void* values, const void* b, void* x, const orc_config* cfg,
                      orc_result* res, double* history, char* err, int errlen);

/* Bounded timing helper for the CPU baseline: runs exactly `iters`
 * preconditioned iterations (no convergence stop) of ICCG (method 0) or
 * COCR (method 1) after setup and returns the wall seconds spent in the
 * iteration loop; *setup_s receives the setup seconds. */
double orc_time_iterations(int dtype, int method, int n, const int* row_ptr,
                           const int* col_idx, const void* values,
                           const void* b, const orc_config* cfg, int iters,
                           double* setup_s);

#ifdef __cplusplus
}
#endif
#endif
