/*
 * C ABI of the MI355X ALS engine: a drop-in for the reference's als_model
 * update path (vserdyuk/cuda_als). Plain pointers and sizes only; no
 * exceptions cross this boundary; every call returns ALS_OK (0) or a
 * negative ALS_E* code (als_status.h) and als_last_error() gives the text.
 *
 * Threading: one engine per host thread. Every device operation is enqueued
 * on the engine's stream (als_set_stream) and calls that return host data
 * synchronise that stream.
 *
 * Reference interface replaced by each entry point (file:line in the
 * reference repository):
 *   als_config_from_reference  INI keys + enums, main.cpp:28-47,
 *                              als_model.h:12-27
 *   als_create / als_set_train / als_set_test
 *                              als_model::als_model, als_model.h:29-32,
 *                              als_model.cu:1517-1552 (copies CSR/CSC/COO to
 *                              the device, cuda_sparse_matrix.cpp:76-110)
 *   als_destroy                als_model::~als_model, als_model.cu:1554-1568
 *   als_init_factors           init in als_model::train, als_model.cu:1772-1784
 *   als_update                 the update-U block (als_model.cu:1797-2029) or
 *                              the update-V block (:2030-2267): RHS
 *                              (csrmm2+geam :1816-1852), Gram (:1891-1982),
 *                              solve (LU_solve_U/V :1570-1720 or
 *                              updateXWithCGHost, cumf_als_cg/cg.h:30-32)
 *   als_sweep / als_train      the sweep loop of als_model::train,
 *                              als_model.cu:1790-2275 (+ CALC_ITER_RMSE :2269)
 *   als_rmse                   als_model::rsme_test / rsme_train,
 *                              als_model.cu:1722-1764
 *   als_get_factors            h_UT / h_VT after train(), als_model.cu:2283-2284
 *   als_save_libmf             als_model::save_libmf, als_model.cu:2294-2312
 *   als_update_rows            LU_solve_U / LU_solve_V (one batch of rows),
 *                              als_model.h:35-36, als_model.cu:1570-1720
 *   als_cg_update_x / als_cg_update_x_fp16 / als_cg_update_x_tt_fp16
 *                              updateXWithCGHost / updateXWithCGHost_tt_fp16,
 *                              cumf_als_cg/cg.h:30-32 (kernels cg.cu:37-232,
 *                              236-430; host side cg.cu:642-688)
 *   als_forward_dev / als_adjoint_dev
 *                              no reference counterpart: differentiable
 *                              half-sweep (adjoint of the batched solve)
 */
#ifndef ALS_H
#define ALS_H

#include <stdint.h>

#include "als_status.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct als_engine als_engine;

/* Solver ids 0..2 keep the reference's SOLVE_TYPE values (als_model.h:23-27).
 * LU and CHOL solve the same SPD system exactly (Cholesky is used for both);
 * CG_FP16 runs the CG recurrence on an A rounded to FP16 storage (what the
 * reference's CUMF_ALS_CG_FP16 intends; its own kernel reinterprets FP32
 * bits as half, cumf_als_cg/cg.cu:642-646, als_model.cu:2003). */
enum { ALS_SOLVE_LU = 0, ALS_SOLVE_CG_FP16 = 1, ALS_SOLVE_CG_FP32 = 2, ALS_SOLVE_CHOL = 3 };
/* Precision the Gram accumulation multiplies in (accumulation is FP32). */
enum { ALS_PREC_FP32 = 0, ALS_PREC_BF16 = 1, ALS_PREC_FP16 = 2 };
enum { ALS_SIDE_USERS = 0, ALS_SIDE_ITEMS = 1 };
enum { ALS_RMSE_TEST = 0, ALS_RMSE_TRAIN = 1 };
enum { ALS_INIT_REFERENCE = 0, ALS_INIT_ZERO = 1 };

typedef struct {
    int f;               /* factor rank, 1..143 */
    float lambda;        /* ALS-WR regularisation, scaled by the row's nnz */
    int iters;           /* sweeps run by als_train when sweeps < 0 */
    int solve_type;      /* ALS_SOLVE_* */
    int gram_prec;       /* ALS_PREC_* */
    int cg_iters;        /* CG iteration cap (reference CG_ITER = 6) */
    float cg_tol;        /* CG exit threshold on r.r (reference CG_ERROR = 1e-4) */
    int refine_steps;    /* FP64-residual refinement steps after Cholesky, 0..3 */
    int m_batches;       /* accepted for config compatibility; the engine never */
    int n_batches;       /*   materialises Gram matrices, so no batching is needed */
    int device;          /* HIP device ordinal */
    int batch_work;      /* ratings per work-queue batch; 0 = chosen by the scheduler */
    int split_len;       /* rows with more ratings are split across waves; 0 = chosen */
    int use_graph;       /* replay each sweep from a captured hipGraph */
    int refine_warm;     /* 1: the first refinement step starts from the row's previous
                            value x0 (the factor before this half-sweep): its FP64
                            residual b - A x0 is formed from the Gram before the
                            factorisation and x = x0 + A^-1 (b - A x0); later steps
                            (refine_steps > 1) refine x as usual. 0: every step
                            refines the fresh solve. */
} als_config;

typedef struct {
    int T;               /* 16-column blocks of the augmented system */
    int ld;              /* row stride (floats) of the device factor matrices */
    int grid[2];         /* persistent waves launched per half-sweep (users, items) */
    int lds_bytes;       /* LDS per wave */
    int nbatches[2];     /* work-queue batches per half-sweep */
    int nunits[2];       /* work units (whole rows + parts of split rows) */
    int nsplit_rows[2];  /* rows split across waves */
    int64_t split_len[2];/* ratings per part of a split row */
    int64_t failures;    /* rows whose factorisation broke down (x set to 0) */
    int64_t m, n, nnz, n_test;
    int dual_max;        /* the dual system's largest rating count + 1 (0: no dual kernel) */
    int dual_rows[2];    /* rows of each side the dual system takes */
    int kernel_variant[2]; /* update kernel variant per side (als_set_kernel_variant) */
    int grid_variant1;   /* persistent waves of the one-wave-per-SIMD variant (0: not built) */
    int dual_limit[2];   /* rows with fewer ratings take the dual system, per side
                            (<= dual_max; als_set_dual_limit, autotuned) */
    int gram_gather[2];  /* 16-bit Gram gathers per side: 1 straight to LDS, 0 through
                            registers (als_set_gram_gather, autotuned) */
} als_info;

int als_config_init(als_config *cfg);
/* Maps the reference's INI/constructor parameters (main.cpp:28-47,
 * als_model.h:29-32). CALCULATE_VTVS_TYPE 0..4 (FP32 Gram variants) map to
 * ALS_PREC_FP32, 5..7 (FP16 tensor-core variants) to ALS_PREC_FP16;
 * smem_col_cnt is accepted and unused. */
int als_config_from_reference(int f, float lambda, int iters, int calculate_vtvs_type,
                              int solve_type, int smem_col_cnt, int m_batches, int n_batches,
                              als_config *out);

int als_create(const als_config *cfg, als_engine **out);
void als_destroy(als_engine *e);
const char *als_last_error(const als_engine *e);
/* hip_stream: a hipStream_t (NULL selects the engine's own stream). A
 * half-sweep may also run a kernel on a second stream of the engine, forked
 * from this stream and joined back into it, so work enqueued on this stream
 * is ordered before and after it as usual. */
int als_set_stream(als_engine *e, void *hip_stream);

/* Host arrays in the reference layout (cuda_sparse_matrix.h:29-37); copied.
 * With a 16-bit Gram (ALS_PREC_FP16 / ALS_PREC_BF16) every row's sum of
 * squared ratings, divided by lambda * n where that product is below 1, must
 * stay below 2^31 (the factorisation's FP16 range); ALS_E_INVALID otherwise. */
int als_set_train(als_engine *e, int64_t m, int64_t n, int64_t nnz,
                  const int *csr_ptr, const int *csr_idx, const float *csr_val,
                  const int *csc_ptr, const int *csc_idx, const float *csc_val);
int als_set_test(als_engine *e, int64_t n_test, const int *row, const int *col, const float *val);

int als_init_factors(als_engine *e, int mode);
/* Host factor matrices, rows x f row-major (the reference's UT/VT layout). */
int als_set_factors(als_engine *e, int side, const float *host, int64_t rows, int f);
int als_get_factors(als_engine *e, int side, float *host, int64_t rows, int f);

int als_update(als_engine *e, int side);
/* Rows [row_begin, row_begin + row_count) of one side only, from the
 * current factors of the other: one batch of the reference's update blocks
 * (LU_solve_U / LU_solve_V with their batch size and offset,
 * als_model.h:35-36, als_model.cu:1570-1720, called per batch at
 * :1996-2011). exact_solve != 0 solves by Cholesky whatever the configured
 * solver (the reference's LU_solve_* always factorise); 0 uses it.
 * Synchronous. */
int als_update_rows(als_engine *e, int side, int64_t row_begin, int64_t row_count, int exact_solve);
int als_sweep(als_engine *e);
/* sweeps < 0 runs cfg.iters; rmse_per_sweep (nullable) receives the test
 * RMSE after every sweep. */
int als_train(als_engine *e, int sweeps, double *rmse_per_sweep);
int als_rmse(als_engine *e, int which, double *out);
int als_synchronize(als_engine *e);
int als_get_info(als_engine *e, als_info *out);
int als_save_libmf(als_engine *e, const char *path);

/* Batched CG on caller device buffers: the reference's updateXWithCGHost
 * (cumf_als_cg/cg.h:30) and updateXWithCGHost_tt_fp16 (cg.h:32), whose
 * cgIter argument becomes cg_iters. A is batch x f x f (symmetric positive
 * definite; row- or column-major alike), x (warm start in, solution out)
 * and b are batch x f, all densely packed device pointers of the current
 * device. Recurrence of the reference kernel (cg.cu:37-232): r = b - A x,
 * p = r; per iteration alpha = r.r / p.A p, x += alpha p, r -= alpha A p,
 * exit once r.r < tol, else p = r + beta p; a zero p.A p, or a zero r.r
 * after an update, ends the loop (the reference divides 0 / 0 there).
 * tol takes the place of the reference's compile-time CG_ERROR (1e-4,
 * cg.cu:32).
 * The _fp16 form takes A stored as IEEE half (the reference's own kernel
 * reinterprets FP32 bits as half, als_model.cu:2003; not reproduced).
 * f in 1..143. Enqueued on hip_stream (null: the default stream) and
 * synchronous, like the reference: returns once the stream has run it. */
int als_cg_update_x(const float *A, float *x, const float *b, int batch, int f, int cg_iters, float tol,
                    void *hip_stream);
int als_cg_update_x_fp16(const uint16_t *A, float *x, const float *b, int batch, int f, int cg_iters, float tol,
                         void *hip_stream);
/* updateXWithCGHost_tt_fp16's own argument types (cumf_als_cg/cg.h:32): an
 * FP32 A whose entries are rounded to FP16 as they are loaded, then the
 * same CG (what the reference's kernel intends with its FP16 A). */
int als_cg_update_x_tt_fp16(const float *A, float *x, const float *b, int batch, int f, int cg_iters, float tol,
                            void *hip_stream);

/* Host-only (no device needed): the work schedule the engine builds for one
 * side from its CSR row pointers: whole rows plus parts of rows longer than
 * split_len, heaviest first, cut into queue batches. units_out (nullable)
 * receives 4 ints per unit {row, kb, ke, split_id}, at most max_units.
 * No reference counterpart (the reference launches one block per row in row
 * order, als_model.cu:1891-1982). */
typedef struct {
    int64_t nunits, nsplit_rows, nbatches, nslots, split_len, batch_work;
} als_schedule_stats;
int als_schedule_preview(int64_t rows, const int *ptr, int split_len, int batch_work, int grid,
                         int gram_prec, als_schedule_stats *out, int *units_out, int64_t max_units);

/* Autotuner (no reference counterpart: the reference launches one block per
 * row with a fixed shape). For each side selected in `sides` (bit 0 users,
 * bit 1 items) times one half-sweep per candidate schedule, split length x
 * batch size around the cost-model defaults (under each kernel variant
 * built), then the update kernel's grid,
 * the dual-system limit (which rows solve n x n instead of f x f systems),
 * (with refinement) the residual's restage threshold and the kernel variant,
 * on the engine's data, and keeps the fastest. Factors and the failure count
 * are left as they were.
 * Call after als_set_train; results are identical in exact arithmetic for
 * every schedule (split rows sum their parts in a fixed order). */
typedef struct {
    int64_t split_len[2], batch_work[2]; /* chosen, per side */
    float default_ms[2], best_ms[2];     /* one half-sweep: default schedule, chosen one */
    int candidates[2];
    int restage_max[2];                  /* refinement: rows up to this long restage their
                                            ratings for the residual, longer ones keep a
                                            copy of the tiles (tuned with refine_steps > 0) */
    int kernel_variant[2];               /* update kernel variant chosen (0: default, 1: one
                                            wave per SIMD; tried where it is built) */
    int grid[2];                         /* persistent waves of the default update kernel
                                            chosen (the occupancy limit or down to half of
                                            it, in steps of one wave per CU; als_set_grid) */
    int dual_limit[2];                   /* rating count below which rows take the dual
                                            system (dual_max or one or two tiles fewer;
                                            als_set_dual_limit) */
    int gram_gather[2];                  /* Gram gathers chosen (als_set_gram_gather) */
} als_tune_result;
int als_autotune(als_engine *e, int sides, als_tune_result *out);
/* Applies one side's schedule directly (e.g. a stored als_autotune result):
 * split length, batch size (<= 0: keep) and the refinement restage threshold
 * (< 0: keep). No reference counterpart. */
int als_set_schedule(als_engine *e, int side, int64_t split_len, int64_t batch_work, int restage_max);
/* Update kernel variant of one side: 0 the default (two waves per SIMD for the
 * refinement solvers), 1 one wave per SIMD with the whole register file
 * (warm-started refinement at 7 to 9 tiles; ALS_E_UNSUPPORTED elsewhere). The
 * results are identical; als_autotune times both. No reference counterpart. */
int als_set_kernel_variant(als_engine *e, int side, int variant);
/* Persistent waves of one side's default update kernel, 1 .. als_info.grid's
 * value at als_create (the occupancy limit); the schedule is rebuilt for
 * them. Results are identical for every grid. No reference counterpart. */
int als_set_grid(als_engine *e, int side, int grid);
/* Rows of one side with fewer than `limit` ratings (and empty rows) solve
 * their n x n dual system; 0 sends every row through the f x f system. 0 ..
 * als_info.dual_max. Both systems have the same solution in exact
 * arithmetic. No reference counterpart. */
int als_set_dual_limit(als_engine *e, int side, int limit);
/* How one side's 16-bit Gram gathers the fixed factors' rows: 1 straight
 * into LDS (LDS-DMA loads, two 32-rating chunks in flight per wave; built for
 * 1-3 and 5-7 tiles), 0 through registers (one chunk in flight; every tile
 * count). Results are identical; als_autotune times both. No reference
 * counterpart. */
int als_set_gram_gather(als_engine *e, int side, int mode);

/* Model groups (no reference counterpart; the reference trains one model per
 * run): up to 4 engines over the same rating matrix (same m, n, nnz and
 * device; any f, precision and solver) updated by ONE graph launch per
 * half-sweep: each model's own kernels, forked onto a stream per model so
 * they run concurrently, joined back into the stream of the engine with
 * the largest f. The launch is ordered after each engine's pending work and
 * before its later work. The graph is captured on first use and again
 * whenever a schedule or buffer it holds changes. Each model's result
 * equals its own als_update exactly (the same kernels on the same work).
 * The engines stay owned by the caller and must outlive the group. */
typedef struct als_group als_group;
int als_group_create(als_engine *const *engines, int count, als_group **out);
void als_group_destroy(als_group *g);
const char *als_group_last_error(const als_group *g);
int als_group_update(als_group *g, int side);
int als_group_sweep(als_group *g);

/* Device factor matrices of the engine (row stride als_info.ld floats). */
int als_device_factors(als_engine *e, int side, float **d_ptr);

/* Differentiable half-sweep on caller device buffers (row stride ld):
 *   als_forward_dev: X = solve(A(lambda_rows, Y), b(Y)) for every row of
 *     `side`; with save_factor the Cholesky factors are kept for the adjoint.
 *   als_adjoint_dev: given gbar = dL/dX, writes xbar = A^-1 gbar,
 *     lambda_bar = -n_u xbar_u . x_u, and ybar (the fixed-factor gradient
 *     through A and b; d_ybar rows x ld is overwritten).
 * Requires solve_type LU or CHOL. d_lambda_rows may be NULL (scalar lambda).
 * The engine keeps ONE saved factorisation (the last saving call's);
 * als_adjoint_dev refuses (ALS_E_STATE) when it belongs to the other side.
 * Callers that chain several half-sweeps in one differentiated graph keep
 * one factorisation per call in a buffer of their own (the _ws forms). */
int als_forward_dev(als_engine *e, int side, const float *d_Y, float *d_X,
                    const float *d_lambda_rows, int save_factor);
int als_adjoint_dev(als_engine *e, int side, const float *d_Y, const float *d_X,
                    const float *d_gbar, float *d_xbar, float *d_lambda_bar, float *d_ybar);
/* Caller-owned factorisation: d_save (device, at least
 * als_factor_save_bytes(side) bytes) receives the forward pass's factor
 * tiles and is read by the matching adjoint. No reference counterpart. */
int als_factor_save_bytes(als_engine *e, int side, int64_t *bytes);
int als_forward_dev_ws(als_engine *e, int side, const float *d_Y, float *d_X,
                       const float *d_lambda_rows, float *d_save, int64_t save_bytes);
int als_adjoint_dev_ws(als_engine *e, int side, const float *d_Y, const float *d_X,
                       const float *d_save, int64_t save_bytes, const float *d_gbar,
                       float *d_xbar, float *d_lambda_bar, float *d_ybar);

/* RMSE over caller device buffers with no allocation. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops). The kernels of als_rmse (FP32
 * dot, FP64 sum, fixed-order reduction, so the same inputs give the same
 * bits) on a COO of n ratings (d_row, d_col, d_val) and factor rows of
 * stride ld floats (16-byte aligned, ld a multiple of 4, f <= ld; columns
 * f.. are ignored): d_out[0] = sum of squared errors, d_out[1] =
 * sqrt(sum / n) (device memory, 2 doubles). The per-block partial sums live
 * in a fixed array of the library on each GPU (1024 doubles, 8 KiB, held
 * until the process ends): n may be anything up to 2^31 - 1 (the engine's
 * rating limit; larger is refused, ALS_E_INVALID). Calls that overlap on one
 * GPU race on that array; order them on one stream. Enqueued on hip_stream
 * (NULL: the default stream); returns without waiting.
 * als_noalloc_last_error() gives the text of this thread's last failure. */
int als_rmse_dev_noalloc(const int *d_row, const int *d_col, const float *d_val, int64_t n,
                         const float *d_U, const float *d_V, int f, int ld, double *d_out,
                         void *hip_stream);
const char *als_noalloc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
