// C++ mirror of the reference's als_model / host_sparse_matrix on top of the
// engine's C ABI.
#include "als_model.hpp"

#include <cerrno>
#include <cstring>
#include <fstream>

#include "als_data.h"

namespace als_mi355x {

namespace {

// Reference: host_sparse_matrix::malloc_load_file, cuda_sparse_matrix.cpp:6-27
// (missing file, empty file and size mismatch are errors).
template <class T>
void load_file(const std::string &path, std::vector<T> &dst, size_t count) {
    std::ifstream ifs(path, std::ios::binary | std::ios::ate);
    if (!ifs) throw std::runtime_error(path + ": " + std::strerror(errno));
    auto size = (size_t)ifs.tellg();
    if (size == 0) throw std::runtime_error(path + " is empty");
    if (size != count * sizeof(T)) throw std::runtime_error("file size does not match buffer size");
    ifs.seekg(0, std::ios::beg);
    dst.resize(count);
    if (!ifs.read(reinterpret_cast<char *>(dst.data()), size)) throw std::runtime_error(path + ": read failed");
}

}  // namespace

host_sparse_matrix::host_sparse_matrix(size_t r, size_t c, size_t v) : row_cnt(r), col_cnt(c), val_cnt(v) {}

void host_sparse_matrix::load_csr(const std::string &vals, const std::string &row_ptrs, const std::string &col_idxs) {
    if (!has_coo) {
        load_file(vals, csr_coo_vals, val_cnt);
        load_file(col_idxs, csr_coo_col_idxs, val_cnt);
    }
    load_file(row_ptrs, csr_row_ptrs, row_cnt + 1);
    has_csr = true;
}

void host_sparse_matrix::load_csc(const std::string &vals, const std::string &row_idxs, const std::string &col_ptrs) {
    load_file(vals, csc_vals, val_cnt);
    load_file(row_idxs, csc_row_idxs, val_cnt);
    load_file(col_ptrs, csc_col_ptrs, col_cnt + 1);
    has_csc = true;
}

void host_sparse_matrix::load_coo(const std::string &vals, const std::string &row_idxs, const std::string &col_idxs) {
    if (!has_csr) {
        load_file(vals, csr_coo_vals, val_cnt);
        load_file(col_idxs, csr_coo_col_idxs, val_cnt);
    }
    load_file(row_idxs, coo_row_idxs, val_cnt);
    has_coo = true;
}

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
        check(als_set_train(engine, m, n, (int64_t)tr.val_cnt, tr.csr_row_ptrs.data(), tr.csr_coo_col_idxs.data(),
                            tr.csr_coo_vals.data(), tr.csc_col_ptrs.data(), tr.csc_row_idxs.data(),
                            tr.csc_vals.data()),
              "als_set_train");
        check(als_set_test(engine, (int64_t)te.val_cnt, te.coo_row_idxs.data(), te.csr_coo_col_idxs.data(),
                           te.csr_coo_vals.data()),
              "als_set_test");
    } catch (...) {
        als_destroy(engine);
        engine = nullptr;
        throw;
    }
    h_UT.assign((size_t)m * f, 0.f);
    h_VT.assign((size_t)n * f, 0.f);
}

als_model::~als_model() { als_destroy(engine); }

void als_model::train(std::vector<double> *rmse) {
    check(als_init_factors(engine, ALS_INIT_REFERENCE), "als_init_factors");
    if (rmse) rmse->assign(iters, 0.0);
    check(als_train(engine, iters, rmse ? rmse->data() : nullptr), "als_train");
    check(als_get_factors(engine, ALS_SIDE_USERS, h_UT.data(), m, f), "als_get_factors(U)");
    check(als_get_factors(engine, ALS_SIDE_ITEMS, h_VT.data(), n, f), "als_get_factors(V)");
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

}  // namespace als_mi355x
