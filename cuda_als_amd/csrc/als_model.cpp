// C++ mirror of the reference's als_model / host_sparse_matrix and the cuMF
// CG host entry points, on top of the engine's C ABI.
#include "als_model.hpp"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>

#include "als_cg.h"
#include "als_data.h"

// ------------------------------------------------------- host_sparse_matrix
// Reference: host_sparse_matrix::malloc_load_file, cuda_sparse_matrix.cpp:6-27
// (a missing, empty or wrongly sized file is an error; the buffer is owned
// by the matrix). Plain host memory: the engine copies it to the device once.
void host_sparse_matrix::malloc_load_file(std::string path, void **host_ptr, size_t buf_size) {
    std::ifstream ifs(path, std::ios::binary | std::ios::ate);
    if (!ifs) throw std::runtime_error(path + ": " + std::strerror(errno));
    const auto size = (size_t)ifs.tellg();
    if (size == 0) throw std::runtime_error(path + " is empty");
    if (size != buf_size) throw std::runtime_error("file size does not match buffer size");
    ifs.seekg(0, std::ios::beg);
    void *buf = std::malloc(buf_size);
    if (!buf) throw std::runtime_error(path + ": out of host memory");
    if (!ifs.read(static_cast<char *>(buf), (std::streamsize)buf_size)) {
        std::free(buf);
        throw std::runtime_error(path + ": read failed");
    }
    std::free(*host_ptr);  // a reload replaces the previous array
    *host_ptr = buf;
}

host_sparse_matrix::host_sparse_matrix(size_t r, size_t c, size_t v) : row_cnt(r), col_cnt(c), val_cnt(v) {}

host_sparse_matrix::~host_sparse_matrix() {
    for (void *p : {(void *)h_csr_coo_vals, (void *)h_csr_row_ptrs, (void *)h_csr_coo_col_idxs,
                    (void *)h_coo_row_idxs, (void *)h_csc_vals, (void *)h_csc_row_idxs, (void *)h_csc_col_ptrs})
        std::free(p);
}

// CSR and COO share the value and column-index arrays: whichever is loaded
// first reads them (cuda_sparse_matrix.cpp:43-69).
void host_sparse_matrix::load_csr(std::string vals, std::string row_ptrs, std::string col_idxs) {
    if (!has_coo) {
        malloc_load_file(vals, (void **)&h_csr_coo_vals, val_cnt * sizeof(float));
        malloc_load_file(col_idxs, (void **)&h_csr_coo_col_idxs, val_cnt * sizeof(int));
    }
    malloc_load_file(row_ptrs, (void **)&h_csr_row_ptrs, (row_cnt + 1) * sizeof(int));
    has_csr = true;
}

void host_sparse_matrix::load_csc(std::string vals, std::string row_idxs, std::string col_ptrs) {
    malloc_load_file(vals, (void **)&h_csc_vals, val_cnt * sizeof(float));
    malloc_load_file(row_idxs, (void **)&h_csc_row_idxs, val_cnt * sizeof(int));
    malloc_load_file(col_ptrs, (void **)&h_csc_col_ptrs, (col_cnt + 1) * sizeof(int));
    has_csc = true;
}

void host_sparse_matrix::load_coo(std::string vals, std::string row_idxs, std::string col_idxs) {
    if (!has_csr) {
        malloc_load_file(vals, (void **)&h_csr_coo_vals, val_cnt * sizeof(float));
        malloc_load_file(col_idxs, (void **)&h_csr_coo_col_idxs, val_cnt * sizeof(int));
    }
    malloc_load_file(row_idxs, (void **)&h_coo_row_idxs, val_cnt * sizeof(int));
    has_coo = true;
}

// ---------------------------------------------------------------- als_model
void als_model::check(int rc, const char *what) {
    if (rc != ALS_OK)
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " +
                                 (engine ? als_last_error(engine) : ""));
}

als_model::als_model(host_sparse_matrix &tr, host_sparse_matrix &te, int f_, float lambda_, int iters_,
                     CALCULATE_VTVS_TYPE vt, SOLVE_TYPE st, int smem, int mb, int nb) {
    als_config cfg;
    if (als_config_from_reference(f_, lambda_, iters_, (int)vt, (int)st, smem, mb, nb, &cfg) != ALS_OK)
        throw std::runtime_error("invalid als_model configuration");
    calculate_vtvs_type = vt;
    solve_type = st;
    smem_col_cnt = smem;
    setup(tr, te, cfg);
}

als_model::als_model(host_sparse_matrix &tr, host_sparse_matrix &te, const als_config &cfg) { setup(tr, te, cfg); }

void als_model::setup(host_sparse_matrix &tr, host_sparse_matrix &te, const als_config &cfg) {
    if (!tr.has_csr || !tr.has_csc) throw std::runtime_error("train ratings need CSR and CSC");
    if (!te.has_coo) throw std::runtime_error("test ratings need COO");
    m = (int)tr.row_cnt;
    n = (int)tr.col_cnt;
    f = cfg.f;
    lambda = cfg.lambda;
    iters = cfg.iters;
    m_batches = cfg.m_batches;
    n_batches = cfg.n_batches;
    int rc = als_create(&cfg, &engine);
    if (rc != ALS_OK) throw std::runtime_error("als_create failed (" + std::to_string(rc) + ")");
    try {
        check(als_set_train(engine, m, n, (int64_t)tr.val_cnt, tr.h_csr_row_ptrs, tr.h_csr_coo_col_idxs,
                            tr.h_csr_coo_vals, tr.h_csc_col_ptrs, tr.h_csc_row_idxs, tr.h_csc_vals),
              "als_set_train");
        check(als_set_test(engine, (int64_t)te.val_cnt, te.h_coo_row_idxs, te.h_csr_coo_col_idxs,
                           te.h_csr_coo_vals),
              "als_set_test");
        h_UT = static_cast<float *>(std::calloc(std::max<size_t>(1, (size_t)m * f), sizeof(float)));
        h_VT = static_cast<float *>(std::calloc(std::max<size_t>(1, (size_t)n * f), sizeof(float)));
        if (!h_UT || !h_VT) throw std::runtime_error("out of host memory for h_UT / h_VT");
    } catch (...) {
        std::free(h_UT);
        std::free(h_VT);
        h_UT = h_VT = nullptr;
        als_destroy(engine);
        engine = nullptr;
        throw;
    }
}

als_model::~als_model() {
    als_destroy(engine);
    std::free(h_UT);
    std::free(h_VT);
}

void als_model::LU_solve_U(int m_batch_size, int m_batch_offset) {
    check(als_update_rows(engine, ALS_SIDE_USERS, m_batch_offset, m_batch_size, 1), "LU_solve_U");
}

void als_model::LU_solve_V(int n_batch_size, int n_batch_offset) {
    check(als_update_rows(engine, ALS_SIDE_ITEMS, n_batch_offset, n_batch_size, 1), "LU_solve_V");
}

void als_model::train(std::vector<double> *rmse) {
    check(als_init_factors(engine, ALS_INIT_REFERENCE), "als_init_factors");
    if (rmse) rmse->assign(iters, 0.0);
    check(als_train(engine, iters, rmse ? rmse->data() : nullptr), "als_train");
    check(als_get_factors(engine, ALS_SIDE_USERS, h_UT, m, f), "als_get_factors(U)");
    check(als_get_factors(engine, ALS_SIDE_ITEMS, h_VT, n, f), "als_get_factors(V)");
}

void als_model::update_U() { check(als_update(engine, ALS_SIDE_USERS), "als_update(U)"); }
void als_model::update_V() { check(als_update(engine, ALS_SIDE_ITEMS), "als_update(V)"); }

float als_model::rsme_train() {
    double r = 0.0;
    check(als_rmse(engine, ALS_RMSE_TRAIN, &r), "als_rmse(train)");
    return (float)r;
}

float als_model::rsme_test() { return (float)rmse_test_fp64(); }

double als_model::rmse_test_fp64() {
    double r = 0.0;
    check(als_rmse(engine, ALS_RMSE_TEST, &r), "als_rmse(test)");
    return r;
}

void als_model::save_libmf(const char *path) { check(als_save_libmf(engine, path), "als_save_libmf"); }

std::string als_model::to_string(CALCULATE_VTVS_TYPE t) {
    switch (t) {
    case CALCULATE_VTVS_TYPE::SIMPLE: return "SIMPLE";
    case CALCULATE_VTVS_TYPE::SMEM_ROW_MAJOR: return "SMEM_ROW_MAJOR";
    case CALCULATE_VTVS_TYPE::SMEM_COL_MAJOR: return "SMEM_COL_MAJOR";
    case CALCULATE_VTVS_TYPE::SMEM_ROW_MAJOR_NO_CALC: return "SMEM_ROW_MAJOR_NO_CALC";
    case CALCULATE_VTVS_TYPE::SMEM_COL_MAJOR_TWO_THREADS: return "SMEM_COL_MAJOR_TWO_THREADS";
    case CALCULATE_VTVS_TYPE::SMEM_ROW_MAJOR_TENSOR: return "SMEM_ROW_MAJOR_TENSOR";
    case CALCULATE_VTVS_TYPE::SMEM_ROW_MAJOR_TENSOR_SYMMETRIC: return "SMEM_ROW_MAJOR_TENSOR_SYMMETRIC";
    case CALCULATE_VTVS_TYPE::SMEM_ROW_MAJOR_TENSOR_SYMMETRIC_MULT_FRAG: return "SMEM_ROW_MAJOR_TENSOR_SYMMETRIC_MULT_FRAG";
    default: return "UNKNOWN";
    }
}

std::string als_model::to_string(SOLVE_TYPE t) {
    switch (t) {
    case SOLVE_TYPE::LU: return "LU";
    case SOLVE_TYPE::CUMF_ALS_CG_FP16: return "CUMF_ALS_CG_FP16";
    case SOLVE_TYPE::CUMF_ALS_CG_FP32: return "CUMF_ALS_CG_FP32";
    default: return "UNKNOWN";
    }
}

// ---------------------------------------------------- cuMF CG host entries
namespace {

// The reference's loop `for (int iter = 0; iter < cgIter; iter++)` with a
// float cgIter runs ceil(cgIter) times (none for cgIter <= 0 or NaN).
int cg_iterations(float cgIter) {
    if (!(cgIter > 0.f)) return 0;
    return cgIter >= 1e9f ? 1000000000 : (int)std::ceil(cgIter);
}

void exit_on(int rc, const char *what) {
    if (rc == ALS_OK) return;
    std::fprintf(stderr, "%s failed (%d): batch, f (1..143) or a HIP error\n", what, rc);
    std::exit(EXIT_FAILURE);
}

constexpr float CG_ERROR = 1e-4f;  // cumf_als_cg/cg.cu:32

}  // namespace

void updateXWithCGHost(float *A, float *x, float *b, const int batchSize, const int f, const float cgIter) {
    exit_on(als_cg_update_x(A, x, b, batchSize, f, cg_iterations(cgIter), CG_ERROR, nullptr), "updateXWithCGHost");
}

void updateXWithCGHost_tt_fp16(float *A, float *x, float *b, const int batchSize, const int f, const float cgIter) {
    exit_on(als_cg_update_x_tt_fp16(A, x, b, batchSize, f, cg_iterations(cgIter), CG_ERROR, nullptr),
            "updateXWithCGHost_tt_fp16");
}
