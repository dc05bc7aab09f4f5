/*
 * Host-side data plumbing of the MI355X ALS engine: the rating-matrix
 * container in the reference's binary layout, its loader/writer, the CSR ->
 * CSC transpose, and the seeded synthetic workload generator.
 *
 * Layout contract (reference cuda_sparse_matrix.h:8-38): CSR of the
 * users x items train matrix (row_ptr int32[m+1], col_idx int32[nnz],
 * val float32[nnz]), its CSC (= CSR of R^T), and a COO test set
 * (row int32, col int32, val float32) sorted by (row, col).
 *
 * Every function returns 0 on success and a negative ALS_E* code otherwise;
 * als_data_last_error() returns the text of the last failure in this thread.
 */
#ifndef ALS_DATA_H
#define ALS_DATA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct als_dataset als_dataset;

enum {
    ALS_SYNTH_UNIFORM_PAIRS = 0,   /* users uniform, items Zipf, distinct pairs */
    ALS_SYNTH_LOGNORMAL_ROWS = 1   /* row sizes lognormal, items Zipf w/o replacement */
};

typedef struct {
    int64_t m, n;            /* users (rows), items (cols) */
    int64_t total;           /* train + test ratings */
    int64_t n_test;          /* held-out ratings, drawn uniformly from the total */
    int row_model;           /* ALS_SYNTH_* */
    double ln_mu, ln_sigma;  /* lognormal row size parameters */
    int64_t row_min, row_max;/* clip bounds on the row size */
    double zipf_s;           /* item popularity exponent */
    int rank;                /* ground-truth rank; coordinates ~ N(0, rank^-1/2) */
    double noise_sigma;      /* additive Gaussian noise */
    double rating_step;      /* 1.0: integer stars, 0.5: half stars */
    double rating_min, rating_max, rating_offset;
    uint64_t seed;
} als_synth_spec;

/* Seeded generator. Replaces the reference's offline data preparation
 * (data/toy_data/prepare_toy_data.py:73-208, data/netflix/prepare_netflix_data.py:58-109)
 * with synthetic data of the same layout; deterministic for a given spec,
 * independent of the thread count. */
int als_synth_generate(const als_synth_spec *spec, als_dataset **out);

/* Build a dataset from caller arrays (CSR train + COO test); the CSC is
 * derived. Arrays are copied. */
int als_dataset_from_csr(int64_t m, int64_t n, int64_t nnz, const int *ptr,
                         const int *idx, const float *val, int64_t n_test,
                         const int *t_row, const int *t_col, const float *t_val,
                         als_dataset **out);

/* Loader for the reference's .bin files (cuda_sparse_matrix.cpp:6-69; file
 * names main.cpp:68-75): R_train_csr.{data,indptr,indices}.bin,
 * R_train_csc.{data,indices,indptr}.bin, R_test_coo.{data,row,col}.bin.
 * File sizes must match the given counts exactly (cuda_sparse_matrix.cpp:20-21). */
int als_dataset_load_bin(const char *folder, int64_t m, int64_t n,
                         int64_t nnz_train, int64_t nnz_test, als_dataset **out);
/* Writer of the same files (plus R_train_coo.row.bin). */
int als_dataset_save_bin(const als_dataset *ds, const char *folder);

int als_dataset_info(const als_dataset *ds, int64_t *m, int64_t *n,
                     int64_t *nnz_train, int64_t *nnz_test);
int als_dataset_csr(const als_dataset *ds, const int **ptr, const int **idx, const float **val);
int als_dataset_csc(const als_dataset *ds, const int **ptr, const int **idx, const float **val);
int als_dataset_test(const als_dataset *ds, const int **row, const int **col, const float **val);
/* Row indices of the train COO in CSR order (reference R_train_coo.row.bin). */
int als_dataset_train_rows(const als_dataset *ds, const int **row);
void als_dataset_free(als_dataset *ds);

/* CSR (rows x cols) -> CSC, i.e. the CSR of the transpose, with row indices
 * ascending inside every column (the order scipy's tocsc() gives, which the
 * reference's CSC files carry). Output arrays are caller-allocated. */
int als_csr_to_csc(int64_t rows, int64_t cols, int64_t nnz, const int *ptr,
                   const int *idx, const float *val, int *cptr, int *cidx, float *cval);

const char *als_data_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
