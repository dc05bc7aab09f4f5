/*
 * FP64 host oracle for the ALS per-row normal-equation update.
 * TEST INFRASTRUCTURE ONLY (see als_oracle.h): never part of the product path.
 */
#include "als_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* Reference: als_model.cu:1772-1777.
 *   srand(0); for k < n*f: h_VT[k] = 0.2*((float)rand() / (float)RAND_MAX);
 *   for k < m*f: h_UT[k] = 0;
 * The expression is float/float, promoted to double by the 0.2 literal and
 * stored to a float; the glibc generator itself is called, not re-derived. */
void orc_reference_init(int64_t m, int64_t n, int f, double *U, double *V) {
    srand(0);
    for (int64_t k = 0; k < n * (int64_t)f; k++) {
        float v = (float)(0.2 * ((float)rand() / (float)RAND_MAX));
        V[k] = (double)v;
    }
    for (int64_t k = 0; k < m * (int64_t)f; k++) U[k] = 0.0;
}

/* Round to a format with `bits` significant bits (including the hidden one)
 * and minimum normal exponent emin, ties to even (nearbyint in the default
 * rounding mode). Overflow is not reachable for factor values. */
static double round_sig(double x, int bits, int emin) {
    if (x == 0.0 || !isfinite(x)) return x;
    int e;
    frexp(fabs(x), &e);           /* |x| = m * 2^e, m in [0.5, 1) */
    int ex = e - 1;               /* floor(log2 |x|) */
    if (ex < emin) ex = emin;     /* subnormal range: fixed quantum */
    double q = ldexp(1.0, ex - (bits - 1));
    return nearbyint(x / q) * q;
}

double orc_round(double x, int prec) {
    switch (prec) {
    case ORC_PREC_BF16: return round_sig((double)(float)x, 8, -126);
    case ORC_PREC_FP16: return round_sig((double)(float)x, 11, -14);
    case ORC_PREC_FP32: return (double)(float)x;
    default: return x;
    }
}

/* Reference Gram: als_model.cu:91-167 (thread t computes column t of
 * A_u = sum y y^T; the diagonal gets items_cnt * lambda at :156-159).
 * Reference RHS: als_model.cu:1816-1852 (csrmm2 R * V^T, transposed by geam
 * so that b_u is contiguous). Both use the same gathered rows y_i. */
void orc_gram_rhs(int f, int64_t nnz_row, const int *idx, const float *val,
                  const double *Y, double lambda, int prec, double *A, double *b) {
    memset(A, 0, sizeof(double) * (size_t)f * f);
    memset(b, 0, sizeof(double) * (size_t)f);
    double *y = (double *)malloc(sizeof(double) * (size_t)f);
    for (int64_t k = 0; k < nnz_row; k++) {
        const double *yr = Y + (int64_t)idx[k] * f;
        for (int j = 0; j < f; j++) y[j] = orc_round(yr[j], prec);
        double r = (double)val[k];
        for (int i = 0; i < f; i++) {
            double yi = y[i];
            double *Ai = A + (size_t)i * f;
            for (int j = 0; j <= i; j++) Ai[j] += yi * y[j];
            b[i] += r * yi;
        }
    }
    for (int i = 0; i < f; i++)
        for (int j = i + 1; j < f; j++) A[(size_t)i * f + j] = A[(size_t)j * f + i];
    for (int i = 0; i < f; i++) A[(size_t)i * f + i] += (double)nnz_row * lambda;
    free(y);
}

/* Textbook Cholesky A = L L^T (row-major, L overwrites the lower triangle),
 * then forward and back substitution. Equals the reference's unpivoted LU
 * solve (cublasSgetrfBatched with PivotArray=NULL, als_model.cu:1594) on the
 * SPD systems ALS produces. */
int orc_chol_solve(int f, double *A, double *b) {
    for (int k = 0; k < f; k++) {
        double d = A[(size_t)k * f + k];
        for (int j = 0; j < k; j++) d -= A[(size_t)k * f + j] * A[(size_t)k * f + j];
        if (!(d > 0.0)) return -1;
        double l = sqrt(d);
        A[(size_t)k * f + k] = l;
        for (int i = k + 1; i < f; i++) {
            double s = A[(size_t)i * f + k];
            for (int j = 0; j < k; j++) s -= A[(size_t)i * f + j] * A[(size_t)k * f + j];
            A[(size_t)i * f + k] = s / l;
        }
    }
    for (int i = 0; i < f; i++) {               /* L z = b */
        double s = b[i];
        for (int j = 0; j < i; j++) s -= A[(size_t)i * f + j] * b[j];
        b[i] = s / A[(size_t)i * f + i];
    }
    for (int i = f - 1; i >= 0; i--) {          /* L^T x = z */
        double s = b[i];
        for (int j = i + 1; j < f; j++) s -= A[(size_t)j * f + i] * b[j];
        b[i] = s / A[(size_t)i * f + i];
    }
    return 0;
}

/* Doolittle LU without pivoting (the reference's getrf with PivotArray=NULL,
 * als_model.cu:1594, then getrs :1615-1627). */
int orc_lu_nopiv_solve(int f, double *A, double *b) {
    for (int k = 0; k < f; k++) {
        double piv = A[(size_t)k * f + k];
        if (piv == 0.0 || !isfinite(piv)) return -1;
        for (int i = k + 1; i < f; i++) {
            double l = A[(size_t)i * f + k] / piv;
            A[(size_t)i * f + k] = l;
            for (int j = k + 1; j < f; j++) A[(size_t)i * f + j] -= l * A[(size_t)k * f + j];
        }
    }
    for (int i = 0; i < f; i++) {
        double s = b[i];
        for (int j = 0; j < i; j++) s -= A[(size_t)i * f + j] * b[j];
        b[i] = s;
    }
    for (int i = f - 1; i >= 0; i--) {
        double s = b[i];
        for (int j = i + 1; j < f; j++) s -= A[(size_t)i * f + j] * b[j];
        b[i] = s / A[(size_t)i * f + i];
    }
    return 0;
}

/* cuMF CG recurrence, cumf_als_cg/cg.cu:37-232:
 *   x0 = current factor (warm start, :49); r = b - A x0; p = r; rsold = r.r
 *   repeat cgIter times: Ap = A p; alpha = rsold / p.Ap; x += alpha p;
 *   r -= alpha Ap; rsnew = r.r; if rsnew < CG_ERROR break (:196);
 *   p = r + (rsnew/rsold) p; rsold = rsnew.
 * Added guards: p.Ap == 0 stops the loop, and so does r.r == 0 after an
 * update (the reference divides 0/0 in both cases). */
int orc_cg_solve(int f, const double *A, const double *b, double *x, int iters, double tol) {
    double *r = (double *)malloc(sizeof(double) * (size_t)f * 3);
    double *p = r + f, *ap = r + 2 * f;
    double rsold = 0.0;
    for (int i = 0; i < f; i++) {
        double s = 0.0;
        for (int j = 0; j < f; j++) s += A[(size_t)i * f + j] * x[j];
        r[i] = b[i] - s;
        p[i] = r[i];
        rsold += r[i] * r[i];
    }
    for (int it = 0; it < iters; it++) {
        double pap = 0.0;
        for (int i = 0; i < f; i++) {
            double s = 0.0;
            for (int j = 0; j < f; j++) s += A[(size_t)i * f + j] * p[j];
            ap[i] = s;
            pap += p[i] * s;
        }
        if (pap == 0.0) break;
        double alpha = rsold / pap;
        double rsnew = 0.0;
        for (int i = 0; i < f; i++) {
            x[i] += alpha * p[i];
            r[i] -= alpha * ap[i];
            rsnew += r[i] * r[i];
        }
        if (rsnew < tol || rsnew == 0.0) break; /* r.r == 0: beta would be 0/0 */
        double beta = rsnew / rsold;
        rsold = rsnew;
        for (int i = 0; i < f; i++) p[i] = r[i] + beta * p[i];
    }
    free(r);
    return 0;
}

/* The recurrence of orc_cg_solve in FP32 arithmetic (ORC_SOLVE_CG32): A, b
 * and x rounded to float, every product, sum and division in float, sums in
 * index order, the exit test on the float r.r against (float)tol. */
int orc_cg_solve_f32(int f, const double *A, const double *b, double *x, int iters, double tol) {
    float *Af = (float *)malloc(sizeof(float) * (size_t)f * f);
    float *v = (float *)malloc(sizeof(float) * (size_t)f * 4);
    float *xf = v, *r = v + f, *p = v + 2 * f, *ap = v + 3 * f;
    const float tolf = (float)tol;
    for (size_t e = 0; e < (size_t)f * f; e++) Af[e] = (float)A[e];
    float rsold = 0.f;
    for (int i = 0; i < f; i++) xf[i] = (float)x[i];
    for (int i = 0; i < f; i++) {
        float s = 0.f;
        for (int j = 0; j < f; j++) s += Af[(size_t)i * f + j] * xf[j];
        r[i] = (float)b[i] - s;
        p[i] = r[i];
        rsold += r[i] * r[i];
    }
    for (int it = 0; it < iters; it++) {
        float pap = 0.f;
        for (int i = 0; i < f; i++) {
            float s = 0.f;
            for (int j = 0; j < f; j++) s += Af[(size_t)i * f + j] * p[j];
            ap[i] = s;
            pap += p[i] * s;
        }
        if (pap == 0.f) break;
        const float alpha = rsold / pap;
        float rsnew = 0.f;
        for (int i = 0; i < f; i++) {
            xf[i] += alpha * p[i];
            r[i] -= alpha * ap[i];
            rsnew += r[i] * r[i];
        }
        if (rsnew < tolf || rsnew == 0.f) break;
        const float beta = rsnew / rsold;
        rsold = rsnew;
        for (int i = 0; i < f; i++) p[i] = r[i] + beta * p[i];
    }
    for (int i = 0; i < f; i++) x[i] = (double)xf[i];
    free(Af);
    free(v);
    return 0;
}

/* One row of the reference's update block (als_model.cu:1891-2011 for users):
 * Gram + RHS, then the configured solver; x holds the warm start for CG. */
static int update_row(const orc_config *cfg, int64_t n_row, const int *idx,
                      const float *val, const double *Y, double lambda,
                      double *A, double *b, double *x) {
    int f = cfg->f;
    if (n_row == 0) {                       /* empty row: x = 0 */
        for (int j = 0; j < f; j++) x[j] = 0.0;
        return 0;
    }
    orc_gram_rhs(f, n_row, idx, val, Y, lambda, cfg->gram_prec, A, b);
    int rc = 0;
    if (cfg->solver == ORC_SOLVE_CG32) {
        rc = orc_cg_solve_f32(f, A, b, x, cfg->cg_iters, cfg->cg_tol);
    } else if (cfg->solver == ORC_SOLVE_CG || cfg->solver == ORC_SOLVE_CG_FP16A) {
        if (cfg->solver == ORC_SOLVE_CG_FP16A)
            for (size_t e = 0; e < (size_t)f * f; e++) A[e] = orc_round(A[e], ORC_PREC_FP16);
        rc = orc_cg_solve(f, A, b, x, cfg->cg_iters, cfg->cg_tol);
    } else {
        rc = cfg->solver == ORC_SOLVE_LU ? orc_lu_nopiv_solve(f, A, b) : orc_chol_solve(f, A, b);
        if (rc == 0) memcpy(x, b, sizeof(double) * (size_t)f);
    }
    if (cfg->store_fp32)
        for (int j = 0; j < f; j++) x[j] = (double)(float)x[j];
    return rc;
}

int64_t orc_update(const orc_config *cfg, int64_t rows, const int *ptr,
                   const int *idx, const float *val, const double *Y,
                   double *X, const double *lambda_rows) {
    int f = cfg->f;
    int64_t bad = 0;
#pragma omp parallel reduction(+ : bad)
    {
        double *A = (double *)malloc(sizeof(double) * (size_t)f * f);
        double *b = (double *)malloc(sizeof(double) * (size_t)f);
#pragma omp for schedule(dynamic, 64)
        for (int64_t u = 0; u < rows; u++) {
            double lam = lambda_rows ? lambda_rows[u] : cfg->lambda;
            int64_t s = ptr[u], e = ptr[u + 1];
            if (update_row(cfg, e - s, idx + s, val + s, Y, lam, A, b, X + u * f) != 0) bad++;
        }
        free(A);
        free(b);
    }
    return bad;
}

int64_t orc_update_rows(const orc_config *cfg, int64_t nsel, const int64_t *sel,
                        const int *ptr, const int *idx, const float *val,
                        const double *Y, double *Xsel_out, const double *Xsel_in,
                        const double *lambda_rows) {
    int f = cfg->f;
    int64_t bad = 0;
#pragma omp parallel reduction(+ : bad)
    {
        double *A = (double *)malloc(sizeof(double) * (size_t)f * f);
        double *b = (double *)malloc(sizeof(double) * (size_t)f);
#pragma omp for schedule(dynamic, 16)
        for (int64_t t = 0; t < nsel; t++) {
            int64_t u = sel[t];
            double lam = lambda_rows ? lambda_rows[u] : cfg->lambda;
            double *x = Xsel_out + t * f;
            if (Xsel_in) memcpy(x, Xsel_in + t * f, sizeof(double) * (size_t)f);
            else memset(x, 0, sizeof(double) * (size_t)f);
            int64_t s = ptr[u], e = ptr[u + 1];
            if (update_row(cfg, e - s, idx + s, val + s, Y, lam, A, b, x) != 0) bad++;
        }
        free(A);
        free(b);
    }
    return bad;
}

/* Reference: calculate_square_error (als_model.cu:62-89) summed by
 * cublasSasum and finished on the host as sqrt(sum / N) (:1744-1764).
 * Here the dot and the sum are FP64; factors are finite so the reference's
 * NaN early-exit in the dot never triggers. */
double orc_rmse(int64_t N, const int *row, const int *col, const float *val,
                const double *U, const double *V, int f) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t t = 0; t < N; t++) {
        const double *u = U + (int64_t)row[t] * f, *v = V + (int64_t)col[t] * f;
        double d = 0.0;
        for (int j = 0; j < f; j++) d += u[j] * v[j];
        double e = (double)val[t] - d;
        s += e * e;
    }
    return sqrt(s / (double)N);
}

double orc_objective(int f, double lambda, int64_t m, const int *csr_ptr,
                     const int *csr_idx, const float *csr_val, int64_t n,
                     const int *csc_ptr, const double *U, const double *V) {
    double loss = 0.0, reg = 0.0;
#pragma omp parallel for reduction(+ : loss, reg) schedule(dynamic, 64)
    for (int64_t u = 0; u < m; u++) {
        const double *x = U + u * f;
        double nx = 0.0;
        for (int j = 0; j < f; j++) nx += x[j] * x[j];
        reg += (double)(csr_ptr[u + 1] - csr_ptr[u]) * nx;
        for (int64_t k = csr_ptr[u]; k < csr_ptr[u + 1]; k++) {
            const double *y = V + (int64_t)csr_idx[k] * f;
            double d = 0.0;
            for (int j = 0; j < f; j++) d += x[j] * y[j];
            double e = (double)csr_val[k] - d;
            loss += e * e;
        }
    }
#pragma omp parallel for reduction(+ : reg) schedule(static)
    for (int64_t i = 0; i < n; i++) {
        const double *y = V + i * f;
        double ny = 0.0;
        for (int j = 0; j < f; j++) ny += y[j] * y[j];
        reg += (double)(csc_ptr[i + 1] - csc_ptr[i]) * ny;
    }
    return loss + lambda * reg;
}

/* Reference training loop: als_model.cu:1766-2289 — per sweep, update U from
 * V with the CSR arrays (:1797-2029), then V from U with the CSC arrays used
 * as the CSR of R^T (:2030-2267). The test RMSE after every sweep is the
 * CALC_ITER_RMSE report (:2269-2274) extended to the last sweep. */
int64_t orc_train(const orc_config *cfg, int64_t m, int64_t n,
                  const int *csr_ptr, const int *csr_idx, const float *csr_val,
                  const int *csc_ptr, const int *csc_idx, const float *csc_val,
                  int64_t n_test, const int *t_row, const int *t_col, const float *t_val,
                  int sweeps, double *U, double *V, double *rmse_out) {
    int64_t bad = 0;
    for (int s = 0; s < sweeps; s++) {
        bad += orc_update(cfg, m, csr_ptr, csr_idx, csr_val, V, U, NULL);
        bad += orc_update(cfg, n, csc_ptr, csc_idx, csc_val, U, V, NULL);
        if (rmse_out) rmse_out[s] = orc_rmse(n_test, t_row, t_col, t_val, U, V, cfg->f);
    }
    return bad;
}

/* Adjoint of x_u = A_u^-1 b_u with A_u = sum y y^T + lambda_u n_u I and
 * b_u = sum r y. For an upstream gradient g_u = dL/dx_u:
 *   xbar_u = A_u^-1 g_u            (A symmetric)
 *   Abar_u = -xbar_u x_u^T,  bbar_u = xbar_u
 *   dL/dlambda_u = trace(Abar_u n_u I) = -n_u xbar_u . x_u
 *   dL/dy_i += r_ui xbar_u - xbar_u (x_u . y_i) - x_u (xbar_u . y_i)
 * evaluated exactly in FP64 with a fresh Cholesky factorisation. */
void orc_adjoint(int f, int64_t rows, int64_t ncols, const int *ptr,
                 const int *idx, const float *val, const double *Y,
                 const double *lambda_rows, const double *X, const double *Gbar,
                 double *Xbar, double *lambda_bar, double *Ybar) {
    (void)ncols;
    double *xb_all = Xbar ? Xbar : (double *)malloc(sizeof(double) * (size_t)rows * f);
#pragma omp parallel
    {
        double *A = (double *)malloc(sizeof(double) * (size_t)f * f);
        double *b = (double *)malloc(sizeof(double) * (size_t)f);
#pragma omp for schedule(dynamic, 64)
        for (int64_t u = 0; u < rows; u++) {
            double *xb = xb_all + u * f;
            int64_t s = ptr[u], e = ptr[u + 1];
            if (e == s) {
                for (int j = 0; j < f; j++) xb[j] = 0.0;
                if (lambda_bar) lambda_bar[u] = 0.0;
                continue;
            }
            orc_gram_rhs(f, e - s, idx + s, val + s, Y, lambda_rows[u], ORC_PREC_EXACT, A, b);
            memcpy(xb, Gbar + u * f, sizeof(double) * (size_t)f);
            orc_chol_solve(f, A, xb);
            double d = 0.0;
            for (int j = 0; j < f; j++) d += xb[j] * X[u * f + j];
            if (lambda_bar) lambda_bar[u] = -(double)(e - s) * d;
        }
        free(A);
        free(b);
    }
    if (Ybar) {
        for (int64_t u = 0; u < rows; u++) {
            const double *x = X + u * f, *xb = xb_all + u * f;
            for (int64_t k = ptr[u]; k < ptr[u + 1]; k++) {
                const double *y = Y + (int64_t)idx[k] * f;
                double *yb = Ybar + (int64_t)idx[k] * f;
                double xy = 0.0, xby = 0.0;
                for (int j = 0; j < f; j++) { xy += x[j] * y[j]; xby += xb[j] * y[j]; }
                double r = (double)val[k];
                for (int j = 0; j < f; j++) yb[j] += r * xb[j] - xb[j] * xy - x[j] * xby;
            }
        }
    }
    if (!Xbar) free(xb_all);
}
