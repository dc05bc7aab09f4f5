// C++ host mirror of the reference's model type, built only on the C ABI
// (als.h, als_data.h). Same names, argument meaning and call order as
// vserdyuk/cuda_als:
//   host_sparse_matrix   cuda_sparse_matrix.h:8-38 (load_csr/load_csc/load_coo,
//                        size-checked file loads, cuda_sparse_matrix.cpp:6-69)
//   als_model            als_model.h:11-80 (ctor :29-32, train, rsme_train,
//                        rsme_test, save_libmf, h_UT/h_VT after train)
// Errors: the reference throws std::runtime_error for configuration and file
// problems and exits on CUDA errors; this mirror throws std::runtime_error
// for both, with the engine's message.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "als.h"

namespace als_mi355x {

struct host_sparse_matrix {
    host_sparse_matrix(size_t row_cnt, size_t col_cnt, size_t val_cnt);

    void load_csr(const std::string &vals_path, const std::string &row_ptrs_path, const std::string &col_idxs_path);
    void load_csc(const std::string &vals_path, const std::string &row_idxs_path, const std::string &col_ptrs_path);
    void load_coo(const std::string &vals_path, const std::string &row_idxs_path, const std::string &col_idxs_path);

    bool has_coo = false, has_csr = false, has_csc = false;
    size_t row_cnt = 0, col_cnt = 0, val_cnt = 0;

    std::vector<float> csr_coo_vals;   // shared by CSR and COO, as in the reference
    std::vector<int> csr_row_ptrs;
    std::vector<int> csr_coo_col_idxs;
    std::vector<int> coo_row_idxs;
    std::vector<float> csc_vals;
    std::vector<int> csc_row_idxs;
    std::vector<int> csc_col_ptrs;
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

    // Reference init rule, iters sweeps (users then items), factors copied
    // back to h_UT / h_VT. rmse_per_sweep (optional) gets the test RMSE
    // after every sweep.
    void train(std::vector<double> *rmse_per_sweep = nullptr);
    // One half-sweep (the reference's update-U / update-V block).
    void update_U();
    void update_V();

    float rsme_train();
    float rsme_test();
    double rmse_test_fp64();

    void save_libmf(const char *path);

    static std::string to_string(CALCULATE_VTVS_TYPE t);
    static std::string to_string(SOLVE_TYPE t);

    int m = 0, n = 0, f = 0;
    float lambda = 0.f;
    int iters = 0;
    int m_batches = 1, n_batches = 1;
    int smem_col_cnt = 0;
    CALCULATE_VTVS_TYPE calculate_vtvs_type = CALCULATE_VTVS_TYPE::SIMPLE;
    SOLVE_TYPE solve_type = SOLVE_TYPE::LU;

    std::vector<float> h_UT;  // m x f, user factors after train()
    std::vector<float> h_VT;  // n x f, item factors after train()

    als_engine *engine = nullptr;

  private:
    void setup(host_sparse_matrix &train, host_sparse_matrix &test, const als_config &cfg);
    void check(int rc, const char *what);
};

}  // namespace als_mi355x
