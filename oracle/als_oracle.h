/*
 * FP64 host oracle for the ALS per-row normal-equation update.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: the parity tests,
 * the smoke check and the CPU-baseline leg of the benchmark may call it.
 * The product path (libals_mi355x.so) never links, loads or calls it.
 *
 * It restates, in plain C with FP64 arithmetic, what the reference computes
 * in its update-U / update-V blocks (citations are to vserdyuk/cuda_als):
 *   Gram      A_u = sum_{i in Omega(u)} y_i y_i^T + lambda * n_u * I
 *             als_model.cu:91-167 (regularisation term :156-159)
 *   RHS       b_u = sum_{i in Omega(u)} r_ui y_i
 *             als_model.cu:1816-1852 (users), :2037-2089 (items via CSC)
 *   solve     unpivoted LU (als_model.cu:1570-1644 -> cublasSgetrfBatched
 *             with PivotArray=NULL), Cholesky on the same SPD system, or the
 *             cuMF conjugate-gradient recurrence (cumf_als_cg/cg.cu:37-232)
 *   RMSE      sqrt(sum (r - u.v)^2 / N)   als_model.cu:62-89, 1744-1764
 *   init      srand(0); V[k] = 0.2 * rand()/RAND_MAX; U = 0
 *             als_model.cu:1772-1777
 *   order     update all users from V, then all items from U
 *             als_model.cu:1797 (U block), :2030 (V block)
 *
 * Semantic decisions where the reference is undefined:
 *   - an empty row (n_u == 0) gets x_u = 0 (the reference divides by zero);
 *   - CG stops when p^T A p == 0 (the reference computes 0/0 = NaN);
 *   - RMSE over an empty set is NaN (sqrt(0/0), as the reference computes).
 *
 * Pinning: every function is checked by tests/test_oracle.py against closed
 * forms, library routines (numpy/scipy), brute force on tiny inputs and
 * invariants; the input layout is pinned by the reference's toy-data logs
 * (tests/golden/). The reference's own FP32 rounding order (cuBLAS LU,
 * WMMA accumulation order) cannot be pinned here: parity is to tolerance.
 */
#ifndef ALS_ORACLE_H
#define ALS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ORC_SOLVE_CG_FP16A: the CG recurrence on A rounded to FP16 storage (b and
 * the CG vectors stay in full precision): what the reference's
 * CUMF_ALS_CG_FP16 solver intends (cumf_als_cg/cg.cu:236-430). */
/* ORC_SOLVE_CG32: the same CG recurrence carried in FP32 (A, b, x and every
 * vector, product and sum rounded to float; sums in index order): what an
 * FP32 execution of the cuMF kernel computes up to its summation order. Its
 * distance to ORC_SOLVE_CG measures how far FP32 rounding alone moves the
 * truncated (6-iteration) CG iterate. */
enum { ORC_SOLVE_LU = 0, ORC_SOLVE_CG_FP16A = 1, ORC_SOLVE_CG = 2, ORC_SOLVE_CHOL = 3, ORC_SOLVE_CG32 = 4 };
enum { ORC_PREC_EXACT = 0, ORC_PREC_BF16 = 1, ORC_PREC_FP16 = 2, ORC_PREC_FP32 = 3 };

typedef struct {
    int f;            /* factor rank */
    double lambda;    /* ALS-WR regularisation (scaled by n_u) */
    int solver;       /* ORC_SOLVE_* */
    int cg_iters;     /* CG iteration cap (reference: CG_ITER 6, als_model.cu:16) */
    double cg_tol;    /* CG exit threshold on r^T r (cumf_als_cg/cg.cu:32,196) */
    int gram_prec;    /* rounding applied to fixed-factor values before Gram/RHS */
    int store_fp32;   /* round every updated factor to float after the solve */
} orc_config;

/* Reference initialisation (als_model.cu:1772-1777), written to doubles. */
void orc_reference_init(int64_t m, int64_t n, int f, double *U, double *V);

/* Round x to the given storage precision (round to nearest, ties to even). */
double orc_round(double x, int prec);

/* Gram + RHS of one row: A (f x f, row-major, full), b (f). */
void orc_gram_rhs(int f, int64_t nnz_row, const int *idx, const float *val,
                  const double *Y, double lambda, int prec, double *A, double *b);

/* In-place dense solvers on one f x f system. Return 0, or -1 on breakdown. */
int orc_chol_solve(int f, double *A, double *b);          /* b <- A^-1 b */
int orc_lu_nopiv_solve(int f, double *A, double *b);      /* b <- A^-1 b */
int orc_cg_solve(int f, const double *A, const double *b, double *x,
                 int iters, double tol);                   /* x in/out */
int orc_cg_solve_f32(int f, const double *A, const double *b, double *x,
                     int iters, double tol);               /* x in/out, FP32 recurrence */

/* One half-sweep: update every row of X from the fixed factors Y.
 * lambda_rows (nullable) overrides cfg->lambda per row.
 * Returns the number of rows whose solve broke down. */
int64_t orc_update(const orc_config *cfg, int64_t rows, const int *ptr,
                   const int *idx, const float *val, const double *Y,
                   double *X, const double *lambda_rows);

/* Same update restricted to the listed rows (sampled parity checks). */
int64_t orc_update_rows(const orc_config *cfg, int64_t nsel, const int64_t *sel,
                        const int *ptr, const int *idx, const float *val,
                        const double *Y, double *Xsel_out, const double *Xsel_in,
                        const double *lambda_rows);

/* RMSE over a COO set (FP64 dot and sum). */
double orc_rmse(int64_t N, const int *row, const int *col, const float *val,
                const double *U, const double *V, int f);

/* Regularised ALS-WR objective: sum (r-u.v)^2 + lambda (sum n_u|u|^2 + sum n_i|v|^2). */
double orc_objective(int f, double lambda, int64_t m, const int *csr_ptr,
                     const int *csr_idx, const float *csr_val, int64_t n,
                     const int *csc_ptr, const double *U, const double *V);

/* Full training: sweeps x (update U from V, then V from U); rmse_out[s] is the
 * test RMSE after sweep s (nullable). U, V hold the initial factors on entry. */
int64_t orc_train(const orc_config *cfg, int64_t m, int64_t n,
                  const int *csr_ptr, const int *csr_idx, const float *csr_val,
                  const int *csc_ptr, const int *csc_idx, const float *csc_val,
                  int64_t n_test, const int *t_row, const int *t_col, const float *t_val,
                  int sweeps, double *U, double *V, double *rmse_out);

/* Adjoint of one half-sweep X = solve(A(lambda, Y), b(Y)) (no reference
 * counterpart). Given Gbar = dL/dX (rows x f) it writes
 *   Xbar        = A^-1 Gbar                       (rows x f, nullable)
 *   lambda_bar  = -n_u * Xbar_u . X_u             (rows)
 *   Ybar       += sum_u r_ui Xbar_u - Xbar_u (X_u.y_i) - X_u (Xbar_u.y_i)
 * The forward solution X must be the exact solve (Cholesky). */
void orc_adjoint(int f, int64_t rows, int64_t ncols, const int *ptr,
                 const int *idx, const float *val, const double *Y,
                 const double *lambda_rows, const double *X, const double *Gbar,
                 double *Xbar, double *lambda_bar, double *Ybar);

/* Number of OpenMP threads the oracle will use. */
int orc_num_threads(void);
/* Sets it (n > 0); used to time the CPU baseline at one core. */
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
