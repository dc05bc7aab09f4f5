// C++ host mirror of the reference's model and rating-matrix types, built
// only on the C ABI (als.h, als_data.h). Same names, member names, argument
// meaning and call order as vserdyuk/cuda_als, so a caller switches by
// including this header instead of als_model.h / cuda_sparse_matrix.h:
//   host_sparse_matrix   cuda_sparse_matrix.h:8-38: load_csr / load_csc /
//                        load_coo (size-checked file loads,
//                        cuda_sparse_matrix.cpp:6-69) into the owned arrays
//                        h_csr_coo_vals, h_csr_row_ptrs, h_csr_coo_col_idxs,
//                        h_coo_row_idxs, h_csc_vals, h_csc_row_idxs,
//                        h_csc_col_ptrs
//   als_model            als_model.h:11-80: the constructor (:29-32),
//                        LU_solve_U / LU_solve_V (:35-36), train, rsme_train,
//                        rsme_test, save_libmf, to_string, and the host
//                        factors h_UT (m x f) / h_VT (n x f) filled by train()
// Errors: the reference throws std::runtime_error for configuration and file
// problems and exits on CUDA errors; this mirror throws std::runtime_error
// for both, with the engine's message.
#pragma once

#include <string>
#include <vector>

#include "als.h"

struct host_sparse_matrix {
    host_sparse_matrix(size_t row_cnt, size_t col_cnt, size_t val_cnt);
    ~host_sparse_matrix();
    host_sparse_matrix(const host_sparse_matrix &) = delete;
    host_sparse_matrix &operator=(const host_sparse_matrix &) = delete;

    void load_csr(std::string vals_path, std::string row_ptrs_path, std::string col_idxs_path);
    void load_csc(std::string vals_path, std::string row_idxs_path, std::string col_ptrs_path);
    void load_coo(std::string vals_path, std::string row_idxs_path, std::string col_idxs_path);

    // Allocates *host_ptr (buf_size bytes, freed by the owner) and fills it
    // from the file, which must hold exactly buf_size bytes.
    static void malloc_load_file(std::string path, void **host_ptr, size_t buf_size);

    bool has_coo = false;
    bool has_csr = false;
    bool has_csc = false;

    size_t row_cnt = 0;
    size_t col_cnt = 0;
    size_t val_cnt = 0;

    float *h_csr_coo_vals = nullptr;   // shared by CSR and COO, as in the reference
    int *h_csr_row_ptrs = nullptr;
    int *h_csr_coo_col_idxs = nullptr;

    int *h_coo_row_idxs = nullptr;

    float *h_csc_vals = nullptr;
    int *h_csc_row_idxs = nullptr;
    int *h_csc_col_ptrs = nullptr;
};

struct als_model {
    enum class CALCULATE_VTVS_TYPE {
        SIMPLE = 0,
        SMEM_ROW_MAJOR,
        SMEM_COL_MAJOR,
        SMEM_ROW_MAJOR_NO_CALC,
        SMEM_COL_MAJOR_TWO_THREADS,
        SMEM_ROW_MAJOR_TENSOR,
        SMEM_ROW_MAJOR_TENSOR_SYMMETRIC,
        SMEM_ROW_MAJOR_TENSOR_SYMMETRIC_MULT_FRAG
    };
    enum class SOLVE_TYPE { LU = 0, CUMF_ALS_CG_FP16, CUMF_ALS_CG_FP32 };

    als_model(host_sparse_matrix &host_train_ratings, host_sparse_matrix &host_test_ratings, int f, float lambda,
              int iters, CALCULATE_VTVS_TYPE calculate_vtvs_type, SOLVE_TYPE solve_type, int smem_col_cnt,
              int m_batches, int n_batches);
    // Extension: full engine configuration (Gram precision, refinement, ...).
    als_model(host_sparse_matrix &host_train_ratings, host_sparse_matrix &host_test_ratings, const als_config &cfg);
    ~als_model();
    als_model(const als_model &) = delete;
    als_model &operator=(const als_model &) = delete;

    // Solve users [m_batch_offset, m_batch_offset + m_batch_size) from the
    // current item factors (the reference's batched LU solve of one batch,
    // which train() calls after that batch's Gram): exact solve, written to
    // the device factors; h_UT is filled by train().
    void LU_solve_U(int m_batch_size, int m_batch_offset);
    void LU_solve_V(int n_batch_size, int n_batch_offset);

    // Reference init rule, iters sweeps (users then items), factors copied
    // back to h_UT / h_VT. rmse_per_sweep (extension, optional) gets the
    // test RMSE after every sweep.
    void train(std::vector<double> *rmse_per_sweep = nullptr);
    // One half-sweep (the reference's update-U / update-V block).
    void update_U();
    void update_V();

    float rsme_train();
    float rsme_test();
    double rmse_test_fp64();

    void save_libmf(const char *path);

    static std::string to_string(CALCULATE_VTVS_TYPE calculate_vtvs_type);
    static std::string to_string(SOLVE_TYPE solve_type);

    int m = 0;  // number of users
    int n = 0;  // number of items

    int m_batches = 1;
    int n_batches = 1;

    int f = 0;  // number of factors
    float lambda = 0.f;
    int iters = 0;

    float *h_VT = nullptr;  // n x f item factors after train()
    float *h_UT = nullptr;  // m x f user factors after train()

    CALCULATE_VTVS_TYPE calculate_vtvs_type = CALCULATE_VTVS_TYPE::SIMPLE;
    SOLVE_TYPE solve_type = SOLVE_TYPE::LU;

    int smem_col_cnt = 0;

    als_engine *engine = nullptr;

  private:
    void setup(host_sparse_matrix &train, host_sparse_matrix &test, const als_config &cfg);
    void check(int rc, const char *what);
};
