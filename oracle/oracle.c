/*
 * oracle.c -- CPU restatement of the reference ABMC-IC(0) Krylov path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): see oracle.h.
 *
 * This file holds the integer part (ABMC schedule) and the dtype dispatch;
 * the scalar-generic numerics live in oracle_impl.inc, included once for
 * double and once for double _Complex.
 *
 * Pinning (how this restatement is itself checked; tests/test_oracle_*.py):
 *   - ABMC arrays: bit-exact against oracle/_ref (the reference headers
 *     compiled here) on the 7-point and curl-curl generators, plus the
 *     invariants (bijection, contiguous sorted blocks, no two same-colour
 *     blocks coupled, colour count >= max(target, 2)).
 *   - IC(0) factor: bit-exact (real) / 1e-14 (complex) against oracle/_ref,
 *     the closed-form pivots of the 1-D [-1 2 -1] operator, the IC(0)
 *     identity (L D^-1 L^T)_ij = A_ij on the pattern, a dense brute-force
 *     IC(0) on small matrices, and the closed-form auto-shift schedule.
 *   - CG / COCR: iteration counts and residual histories against oracle/_ref,
 *     dense numpy solves on small SPD / complex-symmetric systems.
 *
 * Compiled with -ffp-contract=off so that no multiply-add is fused, exactly
 * as the reference compiles for x86-64 without FMA.
 */
#include "oracle.h"

#include <complex.h>
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <time.h>

void orc_config_default(orc_config* c)
{
    /* core/solver_config.hpp:39-157 */
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
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ===================================================================== */
/* ABMC schedule                                                          */
/* ===================================================================== */

typedef struct abmc {
    int n, nc, nb;
    int* color_offsets;   /* nc+1 */
    int* color_blocks;    /* nb   */
    int* block_offsets;   /* nb+1 */
    int* ordering;        /* n: old -> new */
    int* reverse;         /* n: new -> old */
} abmc;

static void abmc_release(abmc* s)
{
    if (!s) return;
    free(s->color_offsets); free(s->color_blocks); free(s->block_offsets);
    free(s->ordering); free(s->reverse);
    free(s);
}

static int cmp_int(const void* a, const void* b)
{
    int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}

/*
 * ABMCSchedule::build (core/abmc_ordering.hpp:83-121) with its four stages:
 * make_algebraic_blocks (:147-227), build_block_graph (:235-273),
 * color_graph (:288-353) and build_row_ordering (:365-410).
 */
static abmc* abmc_make(int n, const int* rp, const int* ci, int block_size,
                       int target_colors)
{
    abmc* s = (abmc*)calloc(1, sizeof(abmc));
    if (!s) return NULL;
    s->n = n;
    if (n <= 0) return s;                       /* :85-88 -> empty schedule */
    if (block_size < 1) block_size = 1;         /* :91 */
    if (target_colors < 2) target_colors = 2;   /* :92 */

    /* ---- stage 1: BFS aggregation (:147-227) ------------------------- */
    const int max_blocks = (n + block_size - 1) / block_size;   /* :95 */
    int* assign = (int*)malloc(sizeof(int) * n);   /* -1 free, -2 queued */
    int* rows = (int*)malloc(sizeof(int) * n);     /* rows, block-major */
    int* bstart = (int*)malloc(sizeof(int) * (max_blocks + 2));
    int* queue = (int*)malloc(sizeof(int) * n);
    for (int i = 0; i < n; ++i) assign[i] = -1;
    int nb = 0, nrows = 0, next_free = 0;
    for (int blk = 0; blk < max_blocks; ++blk) {
        while (next_free < n && assign[next_free] >= 0) ++next_free;  /* :162 */
        if (next_free >= n) break;
        const int id = nb++;
        bstart[id] = nrows;
        int size = 0, qh = 0, qt = 0;
        assign[next_free] = id;
        rows[nrows++] = next_free; ++size;
        for (int k = rp[next_free]; k < rp[next_free + 1]; ++k) {    /* :176 */
            int j = ci[k];
            if (assign[j] == -1) { queue[qt++] = j; assign[j] = -2; }
        }
        while (qh < qt && size < block_size) {                         /* :185 */
            int row = queue[qh++];
            if (assign[row] >= 0) continue;
            assign[row] = id;
            rows[nrows++] = row; ++size;
            if (size >= block_size) break;                             /* :195 */
            for (int k = rp[row]; k < rp[row + 1]; ++k) {
                int j = ci[k];
                if (assign[j] == -1) { queue[qt++] = j; assign[j] = -2; }
            }
        }
        while (qh < qt) {                                              /* :208 */
            int row = queue[qh++];
            if (assign[row] == -2) assign[row] = -1;
        }
    }
    /* stragglers join the last block (:218-226); rows[] stays block-major
     * because the last block is the last segment. */
    for (int i = 0; i < n; ++i) {
        if (assign[i] < 0) {
            if (nb == 0) { bstart[0] = 0; nb = 1; }
            assign[i] = nb - 1;
            rows[nrows++] = i;
        }
    }
    bstart[nb] = nrows;
    free(queue);

    /* ---- stage 2: block graph, first-encounter order (:235-273) ------ */
    int* last_seen = (int*)malloc(sizeof(int) * nb);
    int* brp = (int*)malloc(sizeof(int) * (nb + 1));
    for (int b = 0; b < nb; ++b) last_seen[b] = -1;
    brp[0] = 0;
    for (int bi = 0; bi < nb; ++bi) {
        int cnt = 0;
        for (int q = bstart[bi]; q < bstart[bi + 1]; ++q) {
            int row = rows[q];
            for (int k = rp[row]; k < rp[row + 1]; ++k) {
                int bj = assign[ci[k]];
                if (bi != bj && last_seen[bj] != bi) { last_seen[bj] = bi; ++cnt; }
            }
        }
        brp[bi + 1] = brp[bi] + cnt;
    }
    int* bci = (int*)malloc(sizeof(int) * (brp[nb] > 0 ? brp[nb] : 1));
    for (int b = 0; b < nb; ++b) last_seen[b] = -1;
    for (int bi = 0; bi < nb; ++bi) {
        int pos = brp[bi];
        for (int q = bstart[bi]; q < bstart[bi + 1]; ++q) {
            int row = rows[q];
            for (int k = rp[row]; k < rp[row + 1]; ++k) {
                int bj = assign[ci[k]];
                if (bi != bj && last_seen[bj] != bi) { last_seen[bj] = bi; bci[pos++] = bj; }
            }
        }
    }
    free(last_seen);

    /* ---- stage 3: greedy colouring (:288-353) ------------------------ */
    int ncol = target_colors;
    for (int i = 0; i < nb; ++i) {                                     /* :297 */
        int lower = 0;
        for (int k = brp[i]; k < brp[i + 1]; ++k) if (bci[k] < i) ++lower;
        if (lower >= ncol) ncol = lower + 1;
    }
    int cap = ncol + 1;
    int* color = (int*)malloc(sizeof(int) * nb);
    int* forbidden = (int*)malloc(sizeof(int) * cap);
    int* ccount = (int*)calloc(cap, sizeof(int));
    for (int c = 0; c < cap; ++c) forbidden[c] = -1;
    for (int i = 0; i < nb; ++i) {
        for (int k = brp[i]; k < brp[i + 1]; ++k) {
            int j = bci[k];
            if (j < i && color[j] >= 0) forbidden[color[j]] = i;
        }
        int chosen = -1;
        for (int c = 0; c < ncol; ++c) if (forbidden[c] != i) { chosen = c; break; }
        if (chosen < 0) {                                               /* :337 */
            chosen = ncol++;
            if (ncol + 1 > cap) {
                int ncap = 2 * cap;
                forbidden = (int*)realloc(forbidden, sizeof(int) * ncap);
                ccount = (int*)realloc(ccount, sizeof(int) * ncap);
                for (int c = cap; c < ncap; ++c) { forbidden[c] = -1; ccount[c] = 0; }
                cap = ncap;
            }
        }
        color[i] = chosen;
        ccount[chosen]++;
    }
    free(forbidden);
    free(brp); free(bci);

    /* drop empty colours, keeping order (:349-352) */
    int* cmap = (int*)malloc(sizeof(int) * ncol);
    int nc = 0;
    for (int c = 0; c < ncol; ++c) cmap[c] = ccount[c] > 0 ? nc++ : -1;

    /* ---- stage 4: colour-major row ordering (:365-410) --------------- */
    s->nc = nc; s->nb = nb;
    s->color_offsets = (int*)calloc(nc + 1, sizeof(int));
    s->color_blocks = (int*)malloc(sizeof(int) * (nb > 0 ? nb : 1));
    s->block_offsets = (int*)malloc(sizeof(int) * (nb + 1));
    s->ordering = (int*)malloc(sizeof(int) * n);
    s->reverse = (int*)malloc(sizeof(int) * n);
    for (int c = 0; c < ncol; ++c) if (cmap[c] >= 0) s->color_offsets[cmap[c] + 1] = ccount[c];
    for (int c = 0; c < nc; ++c) s->color_offsets[c + 1] += s->color_offsets[c];
    /* blocks of colour c in increasing original block id (colour lists are
     * filled in node order, :345) */
    int* fill = (int*)malloc(sizeof(int) * (nc > 0 ? nc : 1));
    int* order = (int*)malloc(sizeof(int) * (nb > 0 ? nb : 1));
    for (int c = 0; c < nc; ++c) fill[c] = s->color_offsets[c];
    for (int b = 0; b < nb; ++b) order[fill[cmap[color[b]]]++] = b;
    int row_idx = 0;
    for (int g = 0; g < nb; ++g) {
        int ob = order[g];
        s->color_blocks[g] = g;                                        /* :391 */
        int* seg = rows + bstart[ob];
        int len = bstart[ob + 1] - bstart[ob];
        qsort(seg, (size_t)len, sizeof(int), cmp_int);                 /* :395 */
        s->block_offsets[g] = row_idx;
        for (int q = 0; q < len; ++q) {
            s->ordering[seg[q]] = row_idx;
            s->reverse[row_idx] = seg[q];
            ++row_idx;
        }
    }
    s->block_offsets[nb] = row_idx;
    free(fill); free(order); free(cmap); free(ccount); free(color);
    free(assign); free(rows); free(bstart);
    return s;
}

void* orc_abmc_build(int n, const int* row_ptr, const int* col_idx,
                     int block_size, int target_colors)
{
    if (n < 0 || (n > 0 && (!row_ptr || !col_idx))) return NULL;
    return abmc_make(n, row_ptr, col_idx, block_size, target_colors);
}

int orc_abmc_count(const void* h, int which)
{
    const abmc* s = (const abmc*)h;
    return which == 0 ? s->nc : s->nb;
}

void orc_abmc_copy(const void* h, int which, int* out)
{
    const abmc* s = (const abmc*)h;
    switch (which) {
    case 0: memcpy(out, s->color_offsets, sizeof(int) * (s->nc + 1)); break;
    case 1: memcpy(out, s->color_blocks, sizeof(int) * s->nb); break;
    case 2: memcpy(out, s->block_offsets, sizeof(int) * (s->nb + 1)); break;
    case 3: memcpy(out, s->ordering, sizeof(int) * s->n); break;
    case 4: memcpy(out, s->reverse, sizeof(int) * s->n); break;
    default: break;
    }
}

void orc_abmc_free(void* h) { abmc_release((abmc*)h); }

static void set_err(char* err, int errlen, const char* msg)
{
    if (err && errlen > 0) { strncpy(err, msg, (size_t)errlen - 1); err[errlen - 1] = 0; }
}

/* ===================================================================== */
/* scalar-generic numerics                                                */
/* ===================================================================== */

#define S        double
#define SFX(x)   x##_d
#define IS_CPLX  0
#define CONJ(a)  (a)
#define ABS(a)   fabs(a)
#define NORM2(a) ((a) * (a))
#define SCAL(d)  ((double)(d))
#include "oracle_impl.inc"
#undef S
#undef SFX
#undef IS_CPLX
#undef CONJ
#undef ABS
#undef NORM2
#undef SCAL

#define S        double _Complex
#define SFX(x)   x##_z
#define IS_CPLX  1
#define CONJ(a)  conj(a)
#define ABS(a)   cabs(a)
#define NORM2(a) (creal(a) * creal(a) + cimag(a) * cimag(a))
#define SCAL(d)  CMPLX((double)(d), 0.0)
#include "oracle_impl.inc"

/* ===================================================================== */
/* dtype dispatch                                                         */
/* ===================================================================== */

typedef struct ic_any { int dtype; void* p; } ic_any;

void* orc_ic_setup(int dtype, int n, const int* row_ptr, const int* col_idx,
                   const void* values, const orc_config* cfg,
                   char* err, int errlen)
{
    void* p = dtype == 0
        ? (void*)ic_setup_d(n, row_ptr, col_idx, (const double*)values, cfg, err, errlen)
        : (void*)ic_setup_z(n, row_ptr, col_idx, (const double _Complex*)values, cfg, err, errlen);
    if (!p) return NULL;
    ic_any* h = (ic_any*)malloc(sizeof(ic_any));
    h->dtype = dtype; h->p = p;
    return h;
}

int orc_ic_count(const void* hv, int which)
{
    const ic_any* h = (const ic_any*)hv;
    return h->dtype == 0 ? ic_count_d(h->p, which) : ic_count_z(h->p, which);
}

double orc_ic_actual_shift(const void* hv)
{
    const ic_any* h = (const ic_any*)hv;
    return h->dtype == 0 ? ((ic_d*)h->p)->actual_shift : ((ic_z*)h->p)->actual_shift;
}

void orc_ic_copy_int(const void* hv, int which, int* out)
{
    const ic_any* h = (const ic_any*)hv;
    if (h->dtype == 0) ic_copy_int_d(h->p, which, out); else ic_copy_int_z(h->p, which, out);
}

void orc_ic_copy_val(const void* hv, int which, void* out)
{
    const ic_any* h = (const ic_any*)hv;
    if (h->dtype == 0) ic_copy_val_d(h->p, which, out); else ic_copy_val_z(h->p, which, out);
}

void orc_ic_apply(const void* hv, const void* x, void* y)
{
    const ic_any* h = (const ic_any*)hv;
    if (h->dtype == 0) ic_apply_d((ic_d*)h->p, (const double*)x, (double*)y);
    else ic_apply_z((ic_z*)h->p, (const double _Complex*)x, (double _Complex*)y);
}

void orc_ic_apply_reordered(const void* hv, const void* x, void* y)
{
    const ic_any* h = (const ic_any*)hv;
    if (h->dtype == 0) ic_apply_reordered_d((ic_d*)h->p, (const double*)x, (double*)y);
    else ic_apply_reordered_z((ic_z*)h->p, (const double _Complex*)x, (double _Complex*)y);
}

void orc_ic_free(void* hv)
{
    ic_any* h = (ic_any*)hv;
    if (!h) return;
    if (h->dtype == 0) ic_free_d((ic_d*)h->p); else ic_free_z((ic_z*)h->p);
    free(h);
}

int orc_solve_iccg(int dtype, int n, const int* row_ptr, const int* col_idx,
                   const void* values, const void* b, void* x,
                   const orc_config* cfg, orc_result* res, double* history,
                   char* err, int errlen)
{
    return dtype == 0
        ? solve_iccg_d(n, row_ptr, col_idx, (const double*)values, (const double*)b,
                       (double*)x, cfg, res, history, err, errlen)
        : solve_iccg_z(n, row_ptr, col_idx, (const double _Complex*)values,
                       (const double _Complex*)b, (double _Complex*)x, cfg, res,
                       history, err, errlen);
}

int orc_solve_cocr(int dtype, int n, const int* row_ptr, const int* col_idx,
                   const void* values, const void* b, void* x,
                   const orc_config* cfg, orc_result* res, double* history,
                   char* err, int errlen)
{
    return dtype == 0
        ? solve_cocr_d(n, row_ptr, col_idx, (const double*)values, (const double*)b,
                       (double*)x, cfg, res, history, err, errlen)
        : solve_cocr_z(n, row_ptr, col_idx, (const double _Complex*)values,
                       (const double _Complex*)b, (double _Complex*)x, cfg, res,
                       history, err, errlen);
}

int orc_solve_cocr_pre(int dtype, int n, const int* row_ptr, const int* col_idx,
                       const void* values, const int* m_row_ptr, const int* m_col_idx,
                       const void* m_values, const void* b, void* x,
                       const orc_config* cfg, orc_result* res, double* history,
                       char* err, int errlen)
{
    return dtype == 0
        ? solve_cocr_pre_d(n, row_ptr, col_idx, (const double*)values, m_row_ptr, m_col_idx,
                           (const double*)m_values, (const double*)b, (double*)x, cfg, res,
                           history, err, errlen)
        : solve_cocr_pre_z(n, row_ptr, col_idx, (const double _Complex*)values, m_row_ptr,
                           m_col_idx, (const double _Complex*)m_values,
                           (const double _Complex*)b, (double _Complex*)x, cfg, res, history,
                           err, errlen);
}

int orc_solve_plain(int dtype, int method_cocr, int n, const int* row_ptr,
                    const int* col_idx, const void* values, const void* b,
                    void* x, const orc_config* cfg, orc_result* res,
                    double* history, char* err, int errlen)
{
    return dtype == 0
        ? solve_plain_d(method_cocr, n, row_ptr, col_idx, (const double*)values,
                        (const double*)b, (double*)x, cfg, res, history, err, errlen)
        : solve_plain_z(method_cocr, n, row_ptr, col_idx,
                        (const double _Complex*)values, (const double _Complex*)b,
                        (double _Complex*)x, cfg, res, history, err, errlen);
}

void orc_sgs_apply(int dtype, int n, const int* row_ptr, const int* col_idx,
                   const void* values, const void* x, void* y)
{
    if (dtype == 0) {
        sgs_d* g = sgs_setup_d(n, row_ptr, col_idx, (const double*)values);
        sgs_apply_d(g, (const double*)x, (double*)y);
        sgs_free_d(g);
    } else {
        sgs_z* g = sgs_setup_z(n, row_ptr, col_idx, (const double _Complex*)values);
        sgs_apply_z(g, (const double _Complex*)x, (double _Complex*)y);
        sgs_free_z(g);
    }
}

int orc_solve_sgs(int dtype, int method_cocr, int n, const int* row_ptr, const int* col_idx,
                  const void* values, const void* b, void* x, const orc_config* cfg,
                  orc_result* res, double* history, char* err, int errlen)
{
    return dtype == 0
        ? solve_sgs_pre_d(method_cocr ? 1 : 0, n, row_ptr, col_idx, (const double*)values,
                          (const double*)b, (double*)x, cfg, res, history, err, errlen)
        : solve_sgs_pre_z(method_cocr ? 1 : 0, n, row_ptr, col_idx,
                          (const double _Complex*)values, (const double _Complex*)b,
                          (double _Complex*)x, cfg, res, history, err, errlen);
}

int orc_solve_sgsmrtr(int dtype, int n, const int* row_ptr, const int* col_idx,
                      const void* values, const void* b, void* x, const orc_config* cfg,
                      orc_result* res, double* history, char* err, int errlen)
{
    return dtype == 0
        ? solve_sgsmrtr_d(n, row_ptr, col_idx, (const double*)values, (const double*)b,
                          (double*)x, cfg, res, history, err, errlen)
        : solve_sgsmrtr_z(n, row_ptr, col_idx, (const double _Complex*)values,
                          (const double _Complex*)b, (double _Complex*)x, cfg, res, history,
                          err, errlen);
}

double orc_time_iterations(int dtype, int method, int n, const int* row_ptr,
                           const int* col_idx, const void* values,
                           const void* b, const orc_config* cfg, int iters,
                           double* setup_s)
{
    return dtype == 0
        ? time_iters_d(method, n, row_ptr, col_idx, (const double*)values,
                       (const double*)b, cfg, iters, setup_s)
        : time_iters_z(method, n, row_ptr, col_idx, (const double _Complex*)values,
                       (const double _Complex*)b, cfg, iters, setup_s);
}
