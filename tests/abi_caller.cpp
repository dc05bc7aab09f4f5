// A caller written against the reference's own C++ names, as its main.cpp
// uses them (main.cpp:66-113): host_sparse_matrix with its h_-prefixed
// arrays, the als_model constructor with the reference's enums, train(),
// rsme_test(), rsme_train(), h_UT / h_VT, save_libmf(), LU_solve_U/V with a
// batch size and offset, and the cuMF CG entry updateXWithCGHost
// (cumf_als_cg/cg.h:30). Only the includes differ from a reference caller.
//
//   abi_caller <data_folder> <m> <n> <nnz_train> <nnz_test> <f> <lambda> <iters>
//              <vtvs_type> <solve_type> <model_out> <users_out>
//
// Prints key=value lines; writes the libmf model after train() to
// model_out and, to users_out, the user factors after a re-solve of all
// users from the trained item factors in two LU_solve_U batches.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "als_cg.h"
#include "als_model.hpp"

static void hip_ok(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(2);
    }
}

// x = A^-1 b for one dense SPD system, Gaussian elimination in double.
static std::vector<double> dense_solve(std::vector<double> A, std::vector<double> b, int f) {
    for (int k = 0; k < f; k++)
        for (int i = k + 1; i < f; i++) {
            const double l = A[(size_t)i * f + k] / A[(size_t)k * f + k];
            for (int j = k; j < f; j++) A[(size_t)i * f + j] -= l * A[(size_t)k * f + j];
            b[i] -= l * b[k];
        }
    for (int i = f - 1; i >= 0; i--) {
        double s = b[i];
        for (int j = i + 1; j < f; j++) s -= A[(size_t)i * f + j] * b[j];
        b[i] = s / A[(size_t)i * f + i];
    }
    return b;
}

int main(int argc, char **argv) {
    if (argc != 13) {
        std::fprintf(stderr, "usage: %s data m n nnz_train nnz_test f lambda iters vtvs solve model_out users_out\n",
                     argv[0]);
        return 1;
    }
    const std::string data_folder = argv[1];
    const int m = std::atoi(argv[2]), n = std::atoi(argv[3]);
    const long nnz_train = std::atol(argv[4]), nnz_test = std::atol(argv[5]);
    const int f = std::atoi(argv[6]);
    const float lambda = (float)std::atof(argv[7]);
    const int als_iters = std::atoi(argv[8]);
    const auto vtvs = static_cast<als_model::CALCULATE_VTVS_TYPE>(std::atoi(argv[9]));
    const auto solve = static_cast<als_model::SOLVE_TYPE>(std::atoi(argv[10]));

    host_sparse_matrix host_train_ratings(m, n, nnz_train);
    host_train_ratings.load_csc(data_folder + "/R_train_csc.data.bin", data_folder + "/R_train_csc.indices.bin",
                                data_folder + "/R_train_csc.indptr.bin");
    host_train_ratings.load_csr(data_folder + "/R_train_csr.data.bin", data_folder + "/R_train_csr.indptr.bin",
                                data_folder + "/R_train_csr.indices.bin");
    host_train_ratings.load_coo(data_folder + "/R_train_csr.data.bin", data_folder + "/R_train_coo.row.bin",
                                data_folder + "/R_train_csr.indices.bin");
    host_sparse_matrix host_test_ratings(m, n, nnz_test);
    host_test_ratings.load_coo(data_folder + "/R_test_coo.data.bin", data_folder + "/R_test_coo.row.bin",
                               data_folder + "/R_test_coo.col.bin");
    std::printf("first_rating=%g first_col=%d\n", host_train_ratings.h_csr_coo_vals[0],
                host_train_ratings.h_csr_coo_col_idxs[0]);

    {
        als_model model(host_train_ratings, host_test_ratings, f, lambda, als_iters, vtvs, solve, 32, 1, 1);
        std::printf("solver=%s gram=%s\n", als_model::to_string(model.solve_type).c_str(),
                    als_model::to_string(model.calculate_vtvs_type).c_str());
        model.train();
        std::printf("rmse_test=%.9f\nrmse_train=%.9f\n", model.rsme_test(), model.rsme_train());
        std::printf("h_UT0=%.9g h_VT0=%.9g\n", model.h_UT[0], model.h_VT[0]);
        model.save_libmf(argv[11]);
        // every user again from the trained item factors, in two batches
        const int half = model.m / 2;
        model.LU_solve_U(half, 0);
        model.LU_solve_U(model.m - half, half);
        std::vector<float> users((size_t)model.m * model.f);
        if (als_get_factors(model.engine, ALS_SIDE_USERS, users.data(), model.m, model.f) != ALS_OK) return 3;
        std::ofstream(argv[12], std::ios::binary)
            .write(reinterpret_cast<const char *>(users.data()), (std::streamsize)(users.size() * sizeof(float)));
    }

    // updateXWithCGHost on dense SPD systems A = M M^T + f I, from x = 0,
    // with as many iterations as unknowns: the exact solution up to rounding.
    const int batch = 8, cf = 20;
    std::mt19937 gen(7);
    std::uniform_real_distribution<float> U(-1.f, 1.f);
    std::vector<float> A((size_t)batch * cf * cf), b((size_t)batch * cf), x((size_t)batch * cf, 0.f);
    for (int s = 0; s < batch; s++) {
        std::vector<float> M((size_t)cf * cf);
        for (float &v : M) v = U(gen);
        for (int i = 0; i < cf; i++)
            for (int j = 0; j < cf; j++) {
                double acc = i == j ? cf : 0.0;
                for (int k = 0; k < cf; k++) acc += (double)M[(size_t)i * cf + k] * M[(size_t)j * cf + k];
                A[(size_t)s * cf * cf + (size_t)i * cf + j] = (float)acc;
            }
        for (int i = 0; i < cf; i++) b[(size_t)s * cf + i] = U(gen);
    }
    float *dA = nullptr, *db = nullptr, *dx = nullptr;
    hip_ok(hipMalloc((void **)&dA, A.size() * sizeof(float)), "hipMalloc A");
    hip_ok(hipMalloc((void **)&db, b.size() * sizeof(float)), "hipMalloc b");
    hip_ok(hipMalloc((void **)&dx, x.size() * sizeof(float)), "hipMalloc x");
    hip_ok(hipMemcpy(dA, A.data(), A.size() * sizeof(float), hipMemcpyHostToDevice), "copy A");
    hip_ok(hipMemcpy(db, b.data(), b.size() * sizeof(float), hipMemcpyHostToDevice), "copy b");
    hip_ok(hipMemcpy(dx, x.data(), x.size() * sizeof(float), hipMemcpyHostToDevice), "copy x");
    updateXWithCGHost(dA, dx, db, batch, cf, (float)cf);
    hip_ok(hipMemcpy(x.data(), dx, x.size() * sizeof(float), hipMemcpyDeviceToHost), "copy x back");
    double err = 0.0;
    for (int s = 0; s < batch; s++) {
        std::vector<double> As(A.begin() + (size_t)s * cf * cf, A.begin() + (size_t)(s + 1) * cf * cf);
        std::vector<double> bs(b.begin() + (size_t)s * cf, b.begin() + (size_t)(s + 1) * cf);
        const std::vector<double> xs = dense_solve(As, bs, cf);
        for (int i = 0; i < cf; i++) err = std::fmax(err, std::fabs(xs[i] - x[(size_t)s * cf + i]));
    }
    std::printf("cg_max_abs_err=%.3e\n", err);
    (void)hipFree(dA);
    (void)hipFree(db);
    (void)hipFree(dx);
    return 0;
}
