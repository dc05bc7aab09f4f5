// Host side of the MI355X ALS engine: owns device buffers, builds the
// per-side work schedule, launches the kernels and implements the C ABI of
// include/als.h.
#include "als.h"
#include "als_launch.hpp"
#include "als_sched.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

using namespace als;

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct Side {
    int64_t rows = 0;           // rows of the matrix this side updates
    int *ptr = nullptr, *idx = nullptr;
    float *val = nullptr;
    int4 *units = nullptr, *splits = nullptr;
    int *batch = nullptr, *counters = nullptr;
    float *partial = nullptr;
    int nbatches = 0, nunits = 0, nsplits = 0;
    int64_t nslots = 0, split_len = 0, batch_work = 0;
    int grid = 0;                  // persistent waves of the default update kernel (variant 0),
                                   // at most the engine's grid_max (autotuned per side)
    int variant = 0;               // update kernel variant (1: one wave per SIMD, where built)
    int gram_dma = 1;              // 16-bit Gram gathers straight to LDS (where built) or not
    int restage_max = 0;           // refinement: restage rows up to this long (else tile copy)
    float *lambda_rows = nullptr;  // per-row lambda (engine-owned copy), nullable
    // Light rows (fewer than dual_limit ratings, and empty rows) taken by the
    // dual kernel: their own schedule, whole rows only. dual_limit is at most
    // the engine's dual_max (the tile count the dual kernel is built for);
    // the autotuner picks it per side.
    int dual_limit = 0;
    int4 *dunits = nullptr;
    int *dbatch = nullptr;
    int dnbatches = 0, dnunits = 0;
};

}  // namespace

struct als_engine {
    als_config cfg{};
    int T = 0, ld = 0, prec = PREC_FP32, solve = SOLVE_CHOL, lds = 0;
    int dual_max = 0, dual_lds = 0, dual_grid = 0;  // dual kernel: rows below dual_max ratings (0: none)
    int grid_v1 = 0;                                  // persistent waves of the variant-1 kernel (0: none)
    int grid_max = 0;                                 // occupancy limit of the default update kernel
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    // The dual kernel runs beside the update kernel on a stream of its own,
    // forked from and joined back into `stream` with these events.
    hipStream_t dual_stream = nullptr;
    hipEvent_t dual_fork = nullptr, dual_join = nullptr;
    bool dual_beside = true;  // false: ALS_DUAL_SEQUENTIAL=1 in the environment at als_create
    std::string err;
    int num_cu = 0;

    int64_t m = 0, n = 0, nnz = 0, n_test = 0;
    Side side[2];                      // 0: users (CSR), 1: items (CSC)
    int *coo_row = nullptr;            // train COO rows (train RMSE)
    int *t_row = nullptr, *t_col = nullptr;
    float *t_val = nullptr;
    float *U = nullptr, *V = nullptr;  // FP32 factors, row stride ld
    uint16_t *Us = nullptr, *Vs = nullptr;  // 16-bit shadows (mixed precision)
    int *queue = nullptr;              // 4 ints: update kernel's queue pair, dual kernel's
    int *fail = nullptr;               // 1 int
    double *partials = nullptr, *dsum = nullptr;
    int rmse_blocks = RMSE_BLOCKS;
    float *Lsave = nullptr;            // saved factorisation of als_forward_dev (adjoint)
    size_t Lsave_bytes = 0;
    int Lsave_side = -1;               // side it belongs to (-1: none)
    float *scratch = nullptr;          // per-wave A copy for refinement
    float *zero32 = nullptr;           // one zero row (16T floats) for unused chunk slots
    uint16_t *zero16 = nullptr;
    uint16_t *tmp16 = nullptr;         // shadow of caller-provided Y (forward_dev)
    size_t tmp16_bytes = 0;
    // the adjoint's dL/dY per side: the other side's rating lists cut into
    // parts (YbarArgs), built at the side's first adjoint call
    int4 *yb_parts[2] = {nullptr, nullptr};
    int *yb_colpart[2] = {nullptr, nullptr};
    float *yb_partial[2] = {nullptr, nullptr};
    int yb_nparts[2] = {0, 0};
    hipGraphExec_t graph = nullptr;
    std::vector<int> hptr[2];          // host copies of the CSR / CSC row pointers (schedules)
    std::vector<void *> allocs;
    bool have_train = false, have_test = false;
};

namespace {

int set_err(als_engine *e, int code, const std::string &msg) {
    if (e) e->err = msg;
    return code;
}

#define HIPCHK(e, expr)                                                                        \
    do {                                                                                       \
        hipError_t _st = (expr);                                                               \
        if (_st != hipSuccess)                                                                 \
            return set_err((e), ALS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_st)); \
    } while (0)

template <class T>
int dev_alloc(als_engine *e, T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    void *q = nullptr;
    HIPCHK(e, hipMalloc(&q, count * sizeof(T)));
    e->allocs.push_back(q);
    *p = static_cast<T *>(q);
    return ALS_OK;
}

void dev_free(als_engine *e, void *p) {
    if (!p) return;
    auto it = std::find(e->allocs.begin(), e->allocs.end(), p);
    if (it != e->allocs.end()) e->allocs.erase(it);
    (void)hipFree(p);
}

int blocks_for_T(int f) { return (f + 1 + 15) / 16; }

int check_csr(als_engine *e, int64_t rows, int64_t cols, int64_t nnz, const int *ptr, const int *idx,
              const char *what) {
    if (ptr[0] != 0 || (int64_t)ptr[rows] != nnz)
        return set_err(e, ALS_E_INVALID, std::string(what) + ": pointers must start at 0 and end at nnz");
    for (int64_t r = 0; r < rows; r++)
        if (ptr[r + 1] < ptr[r]) return set_err(e, ALS_E_INVALID, std::string(what) + ": pointers decrease");
    for (int64_t k = 0; k < nnz; k++)
        if (idx[k] < 0 || idx[k] >= cols) return set_err(e, ALS_E_INVALID, std::string(what) + ": index out of range");
    return ALS_OK;
}

// The 16-bit-Gram paths take the Cholesky's trailing updates as FP16
// products of split operands (high part and remainder), so every entry of
// the factor must stay inside FP16's range (65504). The rating row of the
// augmented factor holds z = L^-1 b with |z|^2 = b^T A^-1 b <= sum r^2
// (primal system) and <= sum r^2 / (lambda n) (dual system of a light row,
// whose matrix is at least lambda n I). Rows whose bound reaches 2^31
// (|z| >= 46341) are refused; the FP32 path has no such limit. The Gram
// rows themselves are bounded by sqrt(A_ii), A_ii = sum y_i^2 + lambda n.
int check_fp16_range(als_engine *e, int64_t rows, const int *ptr, const float *val, const char *what) {
    const double lim = 2147483648.0;
    for (int64_t r = 0; r < rows; r++) {
        double s2 = 0;
        for (int64_t k = ptr[r]; k < ptr[r + 1]; k++) s2 += (double)val[k] * val[k];
        const double ln = (double)e->cfg.lambda * (ptr[r + 1] - ptr[r]);
        if (!(s2 < lim) || (ln > 0 && ln < 1 && !(s2 / ln < lim)))
            return set_err(e, ALS_E_INVALID,
                           std::string(what) + " row " + std::to_string(r) +
                               ": its ratings exceed the 16-bit Gram's range (sum of squares, over lambda n "
                               "where that is below 1, must stay below 2^31); use the FP32 Gram");
    }
    return ALS_OK;
}

// Frees one side's schedule buffers.
void free_schedule(als_engine *e, Side &s) {
    dev_free(e, s.units);
    dev_free(e, s.batch);
    dev_free(e, s.splits);
    dev_free(e, s.counters);
    dev_free(e, s.partial);
    dev_free(e, s.dunits);
    dev_free(e, s.dbatch);
    s.units = s.splits = nullptr;
    s.batch = s.counters = nullptr;
    s.partial = nullptr;
    s.dunits = nullptr;
    s.dbatch = nullptr;
    s.nbatches = s.nunits = s.nsplits = 0;
    s.dnbatches = s.dnunits = 0;
}

// (Re)build one side's work schedule over rows [rb, re) and upload it;
// replaces any previous one.
int build_side(als_engine *e, Side &s, int64_t rows, const int *ptr, int64_t split_len, int64_t batch_work,
               int64_t rb = 0, int64_t re = -1) {
    if (re < 0) re = rows;
    // With a dual kernel, the light rows get a schedule of their own.
    std::vector<int> heavy, light;
    const bool ranged = rb > 0 || re < rows;
    const bool dual = s.dual_limit > 0;
    if (dual || ranged) {
        for (int64_t r = rb; r < re; r++) (dual && ptr[r + 1] - ptr[r] < s.dual_limit ? light : heavy).push_back((int)r);
    }
    Schedule sc = build_schedule(rows, ptr, (int)split_len, (int)batch_work, s.grid, e->prec,
                                 dual || ranged ? &heavy : nullptr);
    Schedule sd;
    if (dual && !light.empty()) sd = build_schedule(rows, ptr, (int64_t)1 << 30, 0, e->dual_grid, e->prec, &light);
    static_assert(sizeof(Unit) == sizeof(int4) && sizeof(Split) == sizeof(int4), "unit layout");
    HIPCHK(e, hipStreamSynchronize(e->stream));
    free_schedule(e, s);
    int rc;
    if (!sd.units.empty()) {
        if ((rc = dev_alloc(e, &s.dunits, sd.units.size())) || (rc = dev_alloc(e, &s.dbatch, sd.batch_begin.size())))
            return rc;
        HIPCHK(e, hipMemcpyAsync(s.dunits, sd.units.data(), sd.units.size() * sizeof(int4), hipMemcpyHostToDevice,
                                 e->stream));
        HIPCHK(e, hipMemcpyAsync(s.dbatch, sd.batch_begin.data(), sd.batch_begin.size() * sizeof(int),
                                 hipMemcpyHostToDevice, e->stream));
        s.dnbatches = (int)sd.batch_begin.size() - 1;
        s.dnunits = (int)sd.units.size();
    }
    if ((rc = dev_alloc(e, &s.units, sc.units.size())) || (rc = dev_alloc(e, &s.batch, sc.batch_begin.size())) ||
        (rc = dev_alloc(e, &s.splits, sc.splits.size())) || (rc = dev_alloc(e, &s.counters, sc.splits.size())) ||
        (rc = dev_alloc(e, &s.partial, (size_t)sc.nslots * (e->T * (e->T + 1) / 2) * 256)))
        return rc;
    if (!sc.units.empty())
        HIPCHK(e, hipMemcpyAsync(s.units, sc.units.data(), sc.units.size() * sizeof(int4), hipMemcpyHostToDevice, e->stream));
    if (!sc.splits.empty())
        HIPCHK(e, hipMemcpyAsync(s.splits, sc.splits.data(), sc.splits.size() * sizeof(int4), hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipMemcpyAsync(s.batch, sc.batch_begin.data(), sc.batch_begin.size() * sizeof(int),
                             hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipMemsetAsync(s.counters, 0, std::max<size_t>(1, sc.splits.size()) * sizeof(int), e->stream));
    s.nbatches = (int)sc.batch_begin.size() - 1;
    s.nunits = (int)sc.units.size();
    s.nsplits = (int)sc.splits.size();
    s.nslots = sc.nslots;
    s.split_len = sc.split_len;
    s.batch_work = sc.batch_work;
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return ALS_OK;
}

int drop_graph(als_engine *e) {
    if (e->graph) {
        (void)hipGraphExecDestroy(e->graph);
        e->graph = nullptr;
    }
    return ALS_OK;
}

// Argument block of one half-sweep. X/Y/shadow pointers and lambda are
// explicit so the same path serves the engine's own factors and caller buffers.
UpdateArgs side_args(als_engine *e, const Side &s, const float *Y, const uint16_t *Ys, float *X, uint16_t *Xs,
                     const float *lambda_rows, float *Lsave) {
    UpdateArgs a{};
    a.ptr = s.ptr;
    a.idx = s.idx;
    a.val = s.val;
    a.Y = Y;
    a.Ys = Ys;
    a.X = X;
    a.Xs = Xs;
    a.lambda_rows = lambda_rows;
    a.lambda = e->cfg.lambda;
    a.f = e->cfg.f;
    a.units = s.units;
    a.splits = s.splits;
    a.counters = s.counters;
    a.partial = s.partial;
    a.batch_begin = s.batch;
    a.nbatches = s.nbatches;
    a.queue = e->queue;
    a.cg_iters = e->cfg.cg_iters;
    a.cg_tol = e->cfg.cg_tol;
    a.cg_fp16_a = e->cfg.solve_type == ALS_SOLVE_CG_FP16;
    a.refine = (e->solve == SOLVE_CHOL_REFINE || e->solve == SOLVE_CHOL_WARM) ? e->cfg.refine_steps : 0;
    a.restage_max = s.restage_max;
    a.refine_warm = e->cfg.refine_warm;
    a.fail_count = e->fail;
    a.Lsave = Lsave;
    a.scratch = e->scratch;
    a.zero32 = e->zero32;
    a.zero16 = e->zero16;
    {
        const int64_t ny = &s == &e->side[0] ? e->n : e->m;  // rows of Y
        const int64_t bytes = ny * e->ld * (int64_t)sizeof(uint16_t);
        a.ys_bytes = bytes <= INT32_MAX ? (int)bytes : 0;
    }
    a.gram_dma = s.gram_dma;
    return a;
}

// The same arguments over the dual schedule's units (whole rows only).
UpdateArgs dual_args(const Side &s, UpdateArgs a) {
    a.units = s.dunits;
    a.batch_begin = s.dbatch;
    a.nbatches = s.dnbatches;
    a.splits = nullptr;
    a.counters = nullptr;
    a.partial = nullptr;
    return a;
}

// Launch one half-sweep: the update kernel over its schedule on st and the
// dual kernel over the light rows beside it on the engine's dual stream
// (forked from st before and joined into st after, also under graph
// capture). The two write disjoint rows of X and pop their own work queues;
// running them together lets the dual kernel's waves fill the SIMDs the
// update kernel's last waves leave idle. With beside = false (the model
// group's capture, which forks its own streams) the dual kernel follows
// the update kernel on st. When the factorisation is saved for the adjoint
// (Lsave), the light rows go through the update kernel too, after it: the
// adjoint needs every row's f x f factor.
int launch_schedule(als_engine *e, const Side &s, const float *Y, const uint16_t *Ys, float *X, uint16_t *Xs,
                    const float *lambda_rows, float *Lsave, hipStream_t st, bool beside, int solve) {
    if (s.rows == 0) return ALS_OK;
    UpdateArgs a = side_args(e, s, Y, Ys, X, Xs, lambda_rows, Lsave);
    if (solve == SOLVE_CHOL) a.refine = 0;
    // the side's kernel variant applies to the engine's own solver
    const int var = solve == e->solve ? s.variant : 0;
    const int ugrid = var ? e->grid_v1 : s.grid;
    if (beside && e->dual_beside && !Lsave && s.nbatches > 0 && s.dnbatches > 0) {
        UpdateArgs d = dual_args(s, a);
        d.queue = e->queue + 2;
        HIPCHK(e, hipEventRecord(e->dual_fork, st));
        HIPCHK(e, hipStreamWaitEvent(e->dual_stream, e->dual_fork, 0));
        HIPCHK(e, launch_update(e->T, e->prec, solve, var, a, ugrid, e->lds, st));
        HIPCHK(e, launch_dual(e->T, e->prec, solve, d, e->dual_grid, e->dual_lds, e->dual_stream));
        HIPCHK(e, hipEventRecord(e->dual_join, e->dual_stream));
        HIPCHK(e, hipStreamWaitEvent(st, e->dual_join, 0));
        return ALS_OK;
    }
    if (s.nbatches > 0) HIPCHK(e, launch_update(e->T, e->prec, solve, var, a, ugrid, e->lds, st));
    if (s.dnbatches > 0) {
        const UpdateArgs d = dual_args(s, a);
        if (Lsave) HIPCHK(e, launch_update(e->T, e->prec, solve, var, d, ugrid, e->lds, st));
        else HIPCHK(e, launch_dual(e->T, e->prec, solve, d, e->dual_grid, e->dual_lds, st));
    }
    return ALS_OK;
}

int launch_side(als_engine *e, int sd, const float *Y, const uint16_t *Ys, float *X, uint16_t *Xs,
                const float *lambda_rows, float *Lsave, hipStream_t st, bool beside = true) {
    return launch_schedule(e, e->side[sd], Y, Ys, X, Xs, lambda_rows, Lsave, st, beside, e->solve);
}

// One half-sweep of the engine's own factors on stream st.
int launch_own_side(als_engine *e, int sd, hipStream_t st, bool beside = true) {
    if (sd == ALS_SIDE_USERS)
        return launch_side(e, 0, e->V, e->Vs, e->U, e->Us, e->side[0].lambda_rows, nullptr, st, beside);
    return launch_side(e, 1, e->U, e->Us, e->V, e->Vs, e->side[1].lambda_rows, nullptr, st, beside);
}

int update_side(als_engine *e, int sd) {
    if (!e->have_train) return set_err(e, ALS_E_STATE, "als_set_train has not been called");
    return launch_own_side(e, sd, e->stream);
}

int sweep_direct(als_engine *e) {
    int rc = update_side(e, ALS_SIDE_USERS);
    if (rc) return rc;
    return update_side(e, ALS_SIDE_ITEMS);
}

int rmse_launch(als_engine *e, int which, double *host_out) {
    RmseArgs a{};
    if (which == ALS_RMSE_TEST) {
        if (!e->have_test) return set_err(e, ALS_E_STATE, "als_set_test has not been called");
        a.row = e->t_row;
        a.col = e->t_col;
        a.val = e->t_val;
        a.n = e->n_test;
    } else {
        if (!e->have_train) return set_err(e, ALS_E_STATE, "als_set_train has not been called");
        a.row = e->coo_row;
        a.col = e->side[0].idx;
        a.val = e->side[0].val;
        a.n = e->nnz;
    }
    if (a.n == 0) {
        *host_out = std::nan("");
        return ALS_OK;
    }
    a.U = e->U;
    a.V = e->V;
    a.f = e->cfg.f;
    a.ld = e->ld;
    a.partials = e->partials;
    HIPCHK(e, launch_rmse(a, e->rmse_blocks, e->dsum, false, e->stream));
    double s = 0.0;
    HIPCHK(e, hipMemcpyAsync(&s, e->dsum, sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    *host_out = std::sqrt(s / (double)a.n);
    return ALS_OK;
}

int refresh_shadow(als_engine *e, int sd) {
    if (e->prec == PREC_FP32) return ALS_OK;
    const int64_t rows = sd == 0 ? e->m : e->n;
    const float *src = sd == 0 ? e->U : e->V;
    uint16_t *dst = sd == 0 ? e->Us : e->Vs;
    HIPCHK(e, launch_to16(e->prec, src, dst, rows * e->ld, e->stream));
    return ALS_OK;
}

}  // namespace

extern "C" {

int als_config_init(als_config *c) {
    if (!c) return ALS_E_INVALID;
    std::memset(c, 0, sizeof(*c));
    c->f = 100;
    c->lambda = 0.05f;
    c->iters = 10;
    c->solve_type = ALS_SOLVE_CHOL;
    c->gram_prec = ALS_PREC_FP32;
    c->cg_iters = 6;
    c->cg_tol = 1e-4f;
    c->refine_steps = 0;
    c->m_batches = 1;
    c->n_batches = 1;
    c->device = 0;
    c->batch_work = 0;
    c->split_len = 0;
    c->use_graph = 0;
    return ALS_OK;
}

int als_config_from_reference(int f, float lambda, int iters, int vtvs, int solve_type, int smem_col_cnt,
                              int m_batches, int n_batches, als_config *out) {
    (void)smem_col_cnt;
    if (!out) return ALS_E_INVALID;
    als_config_init(out);
    if (vtvs < 0 || vtvs > 7 || solve_type < 0 || solve_type > 2) return ALS_E_INVALID;
    out->f = f;
    out->lambda = lambda;
    out->iters = iters;
    out->solve_type = solve_type;
    out->gram_prec = vtvs >= 5 ? ALS_PREC_FP16 : ALS_PREC_FP32;
    out->m_batches = m_batches;
    out->n_batches = n_batches;
    return ALS_OK;
}

int als_create(const als_config *cfg, als_engine **out) {
    if (!cfg || !out) return ALS_E_INVALID;
    *out = nullptr;
    if (cfg->f < 1 || blocks_for_T(cfg->f) > MAX_T) return ALS_E_INVALID;
    if (cfg->gram_prec < ALS_PREC_FP32 || cfg->gram_prec > ALS_PREC_FP16) return ALS_E_INVALID;
    if (cfg->solve_type < ALS_SOLVE_LU || cfg->solve_type > ALS_SOLVE_CHOL) return ALS_E_INVALID;
    if (cfg->refine_steps < 0 || cfg->refine_steps > 3 || cfg->cg_iters < 0) return ALS_E_INVALID;
    if (!(cfg->lambda >= 0.f)) return ALS_E_INVALID;
    auto *e = new als_engine;
    e->cfg = *cfg;
    e->T = blocks_for_T(cfg->f);
    e->ld = 16 * e->T;
    e->prec = cfg->gram_prec;
    e->solve = (cfg->solve_type == ALS_SOLVE_CG_FP32 || cfg->solve_type == ALS_SOLVE_CG_FP16) ? SOLVE_CG
               : cfg->refine_steps == 1 && cfg->refine_warm                               ? SOLVE_CHOL_WARM
               : cfg->refine_steps > 0                                                    ? SOLVE_CHOL_REFINE
                                                                                          : SOLVE_CHOL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || cfg->device < 0 || cfg->device >= ndev) {
        delete e;
        return ALS_E_HIP;
    }
    if (hipSetDevice(cfg->device) != hipSuccess) { delete e; return ALS_E_HIP; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, cfg->device) != hipSuccess) { delete e; return ALS_E_HIP; }
    e->num_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&e->own_stream, hipStreamNonBlocking) != hipSuccess) { delete e; return ALS_E_HIP; }
    e->stream = e->own_stream;
    e->lds = update_lds_bytes(e->T, e->prec, e->solve, cfg->refine_steps);
    const void *k = update_kernel_ptr(e->T, e->prec, e->solve);
    int rc = ALS_OK;
    if (!k || e->lds <= 0 || e->lds > (int)prop.sharedMemPerBlock) {
        if (e->lds > 65536 && k)
            (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, e->lds);
        if (!k || e->lds > 163840) rc = set_err(e, ALS_E_UNSUPPORTED, "kernel variant needs too much LDS");
    }
    if (!rc && e->lds > 65536 &&
        hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, e->lds) != hipSuccess)
        rc = set_err(e, ALS_E_UNSUPPORTED, "cannot raise the LDS limit for this kernel variant");
    int per_cu = 0;
    if (!rc && hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 64, e->lds) != hipSuccess)
        rc = set_err(e, ALS_E_HIP, "occupancy query failed");
    if (!rc) {
        per_cu = std::max(1, per_cu);
        e->side[0].grid = e->side[1].grid = e->grid_max = per_cu * e->num_cu;
        // restaging a row moves 2 FP bytes per rating; a round trip of the
        // tile copy 2 NT KiB: equal at 32 (T + 1) ratings
        e->side[0].restage_max = e->side[1].restage_max = 32 * (e->T + 1);
    }
    if (const void *k1 = rc ? nullptr : update_kernel_ptr(e->T, e->prec, e->solve, 1)) {
        if (e->lds > 65536) (void)hipFuncSetAttribute(k1, hipFuncAttributeMaxDynamicSharedMemorySize, e->lds);
        int pc1 = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc1, k1, 64, e->lds) == hipSuccess && pc1 > 0)
            e->grid_v1 = pc1 * e->num_cu;
    }
    const void *dk = dual_kernel_ptr(e->T, e->prec, e->solve);
    if (!rc && dk) {
        e->dual_lds = dual_lds_bytes(e->T, e->prec);
        int dpc = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&dpc, dk, 64, e->dual_lds) != hipSuccess)
            rc = set_err(e, ALS_E_HIP, "occupancy query failed (dual kernel)");
        e->dual_grid = std::max(1, dpc) * e->num_cu;
        e->dual_max = dual_max_ratings(e->T, e->prec);
        e->side[0].dual_limit = e->side[1].dual_limit = e->dual_max;
    }
    if (!rc && e->solve == SOLVE_CHOL_REFINE)
        rc = dev_alloc(e, &e->scratch, (size_t)e->grid_max * (e->T * (e->T + 1) / 2) * 256);
    // Per-dispatch hardware counters (rocprofv3 --pmc) need the two kernels
    // of a half-sweep one after the other: ALS_DUAL_SEQUENTIAL=1 keeps them so.
    if (const char *v = std::getenv("ALS_DUAL_SEQUENTIAL")) e->dual_beside = std::atoi(v) == 0;
    if (!rc && (hipStreamCreateWithFlags(&e->dual_stream, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&e->dual_fork, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&e->dual_join, hipEventDisableTiming) != hipSuccess))
        rc = set_err(e, ALS_E_HIP, "cannot create the dual kernel's stream");
    if (!rc) rc = dev_alloc(e, &e->queue, 4);
    if (!rc) rc = dev_alloc(e, &e->fail, 1);
    if (!rc) rc = dev_alloc(e, &e->partials, e->rmse_blocks);
    if (!rc) rc = dev_alloc(e, &e->dsum, 1);
    if (!rc) rc = dev_alloc(e, &e->zero32, e->ld);
    if (!rc) rc = dev_alloc(e, &e->zero16, e->ld);
    if (!rc) {
        if (hipMemsetAsync(e->queue, 0, 4 * sizeof(int), e->stream) != hipSuccess ||
            hipMemsetAsync(e->fail, 0, sizeof(int), e->stream) != hipSuccess ||
            hipMemsetAsync(e->zero32, 0, e->ld * sizeof(float), e->stream) != hipSuccess ||
            hipMemsetAsync(e->zero16, 0, e->ld * sizeof(uint16_t), e->stream) != hipSuccess ||
            hipStreamSynchronize(e->stream) != hipSuccess)
            rc = set_err(e, ALS_E_HIP, "initial clear failed");
    }
    if (rc) {
        als_destroy(e);
        return rc;
    }
    *out = e;
    return ALS_OK;
}

void als_destroy(als_engine *e) {
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graph(e);
    for (void *p : e->allocs) (void)hipFree(p);
    if (e->dual_stream) (void)hipStreamSynchronize(e->dual_stream);
    if (e->dual_fork) (void)hipEventDestroy(e->dual_fork);
    if (e->dual_join) (void)hipEventDestroy(e->dual_join);
    if (e->dual_stream) (void)hipStreamDestroy(e->dual_stream);
    if (e->own_stream) (void)hipStreamDestroy(e->own_stream);
    delete e;
}

const char *als_last_error(const als_engine *e) { return e ? e->err.c_str() : "null engine"; }

int als_set_stream(als_engine *e, void *s) {
    if (!e) return ALS_E_INVALID;
    drop_graph(e);
    e->stream = s ? static_cast<hipStream_t>(s) : e->own_stream;
    return ALS_OK;
}

int als_set_train(als_engine *e, int64_t m, int64_t n, int64_t nnz, const int *csr_ptr, const int *csr_idx,
                  const float *csr_val, const int *csc_ptr, const int *csc_idx, const float *csc_val) {
    if (!e) return ALS_E_INVALID;
    if (m < 0 || n < 0 || nnz < 0 || nnz >= (int64_t)INT32_MAX || m >= (int64_t)INT32_MAX || n >= (int64_t)INT32_MAX)
        return set_err(e, ALS_E_INVALID, "sizes must be non-negative and fit int32 indices");
    if (!csr_ptr || !csc_ptr || (nnz > 0 && (!csr_idx || !csr_val || !csc_idx || !csc_val)))
        return set_err(e, ALS_E_INVALID, "null train array");
    int rc = check_csr(e, m, n, nnz, csr_ptr, csr_idx, "CSR");
    if (!rc) rc = check_csr(e, n, m, nnz, csc_ptr, csc_idx, "CSC");
    if (!rc && e->prec != PREC_FP32) rc = check_fp16_range(e, m, csr_ptr, csr_val, "CSR");
    if (!rc && e->prec != PREC_FP32) rc = check_fp16_range(e, n, csc_ptr, csc_val, "CSC");
    if (rc) return rc;
    if (e->have_train) return set_err(e, ALS_E_STATE, "train data already set");
    drop_graph(e);
    e->m = m;
    e->n = n;
    e->nnz = nnz;
    Side &su = e->side[0], &si = e->side[1];
    su.rows = m;
    si.rows = n;
    if ((rc = dev_alloc(e, &su.ptr, m + 1)) || (rc = dev_alloc(e, &su.idx, nnz)) || (rc = dev_alloc(e, &su.val, nnz)) ||
        (rc = dev_alloc(e, &si.ptr, n + 1)) || (rc = dev_alloc(e, &si.idx, nnz)) || (rc = dev_alloc(e, &si.val, nnz)) ||
        (rc = dev_alloc(e, &e->coo_row, nnz)) ||
        (rc = dev_alloc(e, &e->U, m * e->ld)) || (rc = dev_alloc(e, &e->V, n * e->ld)))
        return rc;
    if (e->prec != PREC_FP32 && ((rc = dev_alloc(e, &e->Us, m * e->ld)) || (rc = dev_alloc(e, &e->Vs, n * e->ld))))
        return rc;
    HIPCHK(e, hipMemcpyAsync(su.ptr, csr_ptr, (m + 1) * sizeof(int), hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipMemcpyAsync(si.ptr, csc_ptr, (n + 1) * sizeof(int), hipMemcpyHostToDevice, e->stream));
    if (nnz > 0) {
        HIPCHK(e, hipMemcpyAsync(su.idx, csr_idx, nnz * sizeof(int), hipMemcpyHostToDevice, e->stream));
        HIPCHK(e, hipMemcpyAsync(su.val, csr_val, nnz * sizeof(float), hipMemcpyHostToDevice, e->stream));
        HIPCHK(e, hipMemcpyAsync(si.idx, csc_idx, nnz * sizeof(int), hipMemcpyHostToDevice, e->stream));
        HIPCHK(e, hipMemcpyAsync(si.val, csc_val, nnz * sizeof(float), hipMemcpyHostToDevice, e->stream));
        std::vector<int> rows(nnz);
        for (int64_t r = 0; r < m; r++)
            for (int64_t k = csr_ptr[r]; k < csr_ptr[r + 1]; k++) rows[k] = (int)r;
        HIPCHK(e, hipMemcpyAsync(e->coo_row, rows.data(), nnz * sizeof(int), hipMemcpyHostToDevice, e->stream));
        HIPCHK(e, hipStreamSynchronize(e->stream));
    }
    HIPCHK(e, hipMemsetAsync(e->U, 0, (size_t)std::max<int64_t>(1, m * e->ld) * sizeof(float), e->stream));
    HIPCHK(e, hipMemsetAsync(e->V, 0, (size_t)std::max<int64_t>(1, n * e->ld) * sizeof(float), e->stream));
    if (e->Us) HIPCHK(e, hipMemsetAsync(e->Us, 0, (size_t)std::max<int64_t>(1, m * e->ld) * 2, e->stream));
    if (e->Vs) HIPCHK(e, hipMemsetAsync(e->Vs, 0, (size_t)std::max<int64_t>(1, n * e->ld) * 2, e->stream));
    e->hptr[0].assign(csr_ptr, csr_ptr + m + 1);
    e->hptr[1].assign(csc_ptr, csc_ptr + n + 1);
    for (int sd = 0; sd < 2; ++sd) {  // the adjoint's parts follow the rating lists
        dev_free(e, e->yb_parts[sd]);
        dev_free(e, e->yb_colpart[sd]);
        dev_free(e, e->yb_partial[sd]);
        e->yb_parts[sd] = nullptr;
        e->yb_colpart[sd] = nullptr;
        e->yb_partial[sd] = nullptr;
        e->yb_nparts[sd] = 0;
    }
    if ((rc = build_side(e, su, m, csr_ptr, e->cfg.split_len, e->cfg.batch_work)) ||
        (rc = build_side(e, si, n, csc_ptr, e->cfg.split_len, e->cfg.batch_work)))
        return rc;
    e->have_train = true;
    return ALS_OK;
}

int als_set_test(als_engine *e, int64_t n_test, const int *row, const int *col, const float *val) {
    if (!e) return ALS_E_INVALID;
    if (!e->have_train) return set_err(e, ALS_E_STATE, "set the train data first");
    if (n_test < 0 || n_test >= (int64_t)INT32_MAX) return set_err(e, ALS_E_INVALID, "bad test size");
    if (n_test > 0 && (!row || !col || !val)) return set_err(e, ALS_E_INVALID, "null test array");
    for (int64_t t = 0; t < n_test; t++)
        if (row[t] < 0 || row[t] >= e->m || col[t] < 0 || col[t] >= e->n)
            return set_err(e, ALS_E_INVALID, "test index out of range");
    if (e->t_row) { dev_free(e, e->t_row); dev_free(e, e->t_col); dev_free(e, e->t_val); }
    int rc;
    if ((rc = dev_alloc(e, &e->t_row, n_test)) || (rc = dev_alloc(e, &e->t_col, n_test)) ||
        (rc = dev_alloc(e, &e->t_val, n_test)))
        return rc;
    if (n_test > 0) {
        HIPCHK(e, hipMemcpyAsync(e->t_row, row, n_test * sizeof(int), hipMemcpyHostToDevice, e->stream));
        HIPCHK(e, hipMemcpyAsync(e->t_col, col, n_test * sizeof(int), hipMemcpyHostToDevice, e->stream));
        HIPCHK(e, hipMemcpyAsync(e->t_val, val, n_test * sizeof(float), hipMemcpyHostToDevice, e->stream));
        HIPCHK(e, hipStreamSynchronize(e->stream));
    }
    e->n_test = n_test;
    e->have_test = true;
    return ALS_OK;
}

// Reference: als_model.cu:1772-1777. glibc's srand(0)/rand() sequence is
// reproduced with the reentrant random_r on a private TYPE_3 state (the
// state size srand uses), so the process-wide rand() state is untouched.
int als_init_factors(als_engine *e, int mode) {
    if (!e) return ALS_E_INVALID;
    if (!e->have_train) return set_err(e, ALS_E_STATE, "set the train data first");
    const int f = e->cfg.f;
    std::vector<float> Vh((size_t)e->n * e->ld, 0.f);
    if (mode == ALS_INIT_REFERENCE) {
        struct random_data rd;
        char state[128];
        std::memset(&rd, 0, sizeof(rd));
        std::memset(state, 0, sizeof(state));
        initstate_r(0u, state, sizeof(state), &rd);
        for (int64_t k = 0; k < e->n * f; k++) {
            int32_t r = 0;
            random_r(&rd, &r);
            Vh[(size_t)(k / f) * e->ld + (k % f)] = (float)(0.2 * ((float)r / (float)RAND_MAX));
        }
    } else if (mode != ALS_INIT_ZERO) {
        return set_err(e, ALS_E_INVALID, "unknown init mode");
    }
    HIPCHK(e, hipMemcpyAsync(e->V, Vh.data(), Vh.size() * sizeof(float), hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipMemsetAsync(e->U, 0, (size_t)std::max<int64_t>(1, e->m * e->ld) * sizeof(float), e->stream));
    int rc = refresh_shadow(e, 0);
    if (!rc) rc = refresh_shadow(e, 1);
    if (rc) return rc;
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return ALS_OK;
}

int als_set_factors(als_engine *e, int sd, const float *host, int64_t rows, int f) {
    if (!e || !host) return ALS_E_INVALID;
    if (!e->have_train) return set_err(e, ALS_E_STATE, "set the train data first");
    if (sd != 0 && sd != 1) return set_err(e, ALS_E_INVALID, "bad side");
    if (f != e->cfg.f || rows != (sd == 0 ? e->m : e->n)) return set_err(e, ALS_E_INVALID, "shape mismatch");
    float *dst = sd == 0 ? e->U : e->V;
    HIPCHK(e, hipMemsetAsync(dst, 0, (size_t)std::max<int64_t>(1, rows * e->ld) * sizeof(float), e->stream));
    if (rows > 0)
        HIPCHK(e, hipMemcpy2DAsync(dst, e->ld * sizeof(float), host, f * sizeof(float), f * sizeof(float), rows,
                                   hipMemcpyHostToDevice, e->stream));
    int rc = refresh_shadow(e, sd);
    if (rc) return rc;
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return ALS_OK;
}

int als_get_factors(als_engine *e, int sd, float *host, int64_t rows, int f) {
    if (!e || !host) return ALS_E_INVALID;
    if (!e->have_train) return set_err(e, ALS_E_STATE, "set the train data first");
    if (sd != 0 && sd != 1) return set_err(e, ALS_E_INVALID, "bad side");
    if (f != e->cfg.f || rows != (sd == 0 ? e->m : e->n)) return set_err(e, ALS_E_INVALID, "shape mismatch");
    const float *src = sd == 0 ? e->U : e->V;
    if (rows > 0)
        HIPCHK(e, hipMemcpy2DAsync(host, f * sizeof(float), src, e->ld * sizeof(float), f * sizeof(float), rows,
                                   hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return ALS_OK;
}

int als_device_factors(als_engine *e, int sd, float **d) {
    if (!e || !d || (sd != 0 && sd != 1)) return ALS_E_INVALID;
    *d = sd == 0 ? e->U : e->V;
    return ALS_OK;
}

int als_update(als_engine *e, int sd) {
    if (!e) return ALS_E_INVALID;
    if (sd != 0 && sd != 1) return set_err(e, ALS_E_INVALID, "bad side");
    return update_side(e, sd);
}

// Rows [row_begin, row_begin + row_count) of one side from the current
// factors of the other: a schedule of its own for the range (built, run and
// freed here), the engine's kernels. exact_solve: the plain Cholesky solve
// whatever the engine's solver (the reference's LU_solve_U/V always
// factorises).
int als_update_rows(als_engine *e, int sd, int64_t row_begin, int64_t row_count, int exact_solve) {
    if (!e) return ALS_E_INVALID;
    if (!e->have_train) return set_err(e, ALS_E_STATE, "als_set_train has not been called");
    if (sd != 0 && sd != 1) return set_err(e, ALS_E_INVALID, "bad side");
    const Side &s0 = e->side[sd];
    if (row_begin < 0 || row_count < 0 || row_begin + row_count > s0.rows)
        return set_err(e, ALS_E_INVALID, "row range outside the side's rows");
    if (row_count == 0) return ALS_OK;
    drop_graph(e);
    Side t = s0;  // same arrays, grid and parameters; a schedule of its own
    t.units = t.splits = nullptr;
    t.batch = t.counters = nullptr;
    t.partial = nullptr;
    t.dunits = nullptr;
    t.dbatch = nullptr;
    int rc = build_side(e, t, s0.rows, e->hptr[sd].data(), s0.split_len, s0.batch_work, row_begin,
                        row_begin + row_count);
    const int solve = exact_solve ? SOLVE_CHOL : e->solve;
    if (!rc) {
        rc = sd == ALS_SIDE_USERS
                 ? launch_schedule(e, t, e->V, e->Vs, e->U, e->Us, t.lambda_rows, nullptr, e->stream, true, solve)
                 : launch_schedule(e, t, e->U, e->Us, e->V, e->Vs, t.lambda_rows, nullptr, e->stream, true, solve);
    }
    const hipError_t st = hipStreamSynchronize(e->stream);
    free_schedule(e, t);
    if (!rc && st != hipSuccess) rc = set_err(e, ALS_E_HIP, std::string("update_rows: ") + hipGetErrorString(st));
    return rc;
}

int als_sweep(als_engine *e) {
    if (!e) return ALS_E_INVALID;
    if (!e->have_train) return set_err(e, ALS_E_STATE, "als_set_train has not been called");
    if (!e->cfg.use_graph) return sweep_direct(e);
    if (!e->graph) {
        hipGraph_t g = nullptr;
        HIPCHK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        int rc = sweep_direct(e);
        hipError_t st = hipStreamEndCapture(e->stream, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (st != hipSuccess) return set_err(e, ALS_E_HIP, std::string("graph capture: ") + hipGetErrorString(st));
        st = hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (st != hipSuccess) { e->graph = nullptr; return set_err(e, ALS_E_HIP, "graph instantiate failed"); }
    }
    HIPCHK(e, hipGraphLaunch(e->graph, e->stream));
    return ALS_OK;
}

int als_train(als_engine *e, int sweeps, double *rmse) {
    if (!e) return ALS_E_INVALID;
    if (sweeps < 0) sweeps = e->cfg.iters;
    for (int s = 0; s < sweeps; s++) {
        int rc = als_sweep(e);
        if (rc) return rc;
        if (rmse) {
            rc = als_rmse(e, ALS_RMSE_TEST, &rmse[s]);
            if (rc) return rc;
        }
    }
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return ALS_OK;
}

int als_rmse(als_engine *e, int which, double *out) {
    if (!e || !out) return ALS_E_INVALID;
    if (which != ALS_RMSE_TEST && which != ALS_RMSE_TRAIN) return set_err(e, ALS_E_INVALID, "bad rmse set");
    return rmse_launch(e, which, out);
}

int als_synchronize(als_engine *e) {
    if (!e) return ALS_E_INVALID;
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return ALS_OK;
}

int als_get_info(als_engine *e, als_info *o) {
    if (!e || !o) return ALS_E_INVALID;
    std::memset(o, 0, sizeof(*o));
    o->T = e->T;
    o->ld = e->ld;
    o->grid[0] = e->side[0].grid;
    o->grid[1] = e->side[1].grid;
    o->lds_bytes = e->lds;
    for (int sd = 0; sd < 2; sd++) {
        o->nbatches[sd] = e->side[sd].nbatches + e->side[sd].dnbatches;
        o->nunits[sd] = e->side[sd].nunits + e->side[sd].dnunits;
        o->nsplit_rows[sd] = e->side[sd].nsplits;
        o->split_len[sd] = e->side[sd].split_len;
    }
    int f = 0;
    HIPCHK(e, hipMemcpyAsync(&f, e->fail, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    o->failures = f;
    o->m = e->m;
    o->n = e->n;
    o->nnz = e->nnz;
    o->n_test = e->n_test;
    o->dual_max = e->dual_max;
    o->dual_limit[0] = e->side[0].dual_limit;
    o->dual_limit[1] = e->side[1].dual_limit;
    o->dual_rows[0] = e->side[0].dnunits;
    o->dual_rows[1] = e->side[1].dnunits;
    for (int sd = 0; sd < 2; sd++) {
        o->kernel_variant[sd] = e->side[sd].variant;
        o->gram_gather[sd] = gram_dma_built(e->T, e->prec) ? e->side[sd].gram_dma : 0;
        o->grid_variant1 = e->grid_v1;
    }
    return ALS_OK;
}

// Reference: als_model::save_libmf (als_model.cu:2294-2312), the cumf_sgd /
// libmf binary model: int loss=0, int m, int n, int f, float mean rating,
// then U (m x f) and V (n x f) as float32. The mean is summed in FP64 here
// (the reference uses cublasSasum in FP32).
int als_save_libmf(als_engine *e, const char *path) {
    if (!e || !path) return ALS_E_INVALID;
    if (!e->have_train) return set_err(e, ALS_E_STATE, "set the train data first");
    const int f = e->cfg.f;
    std::vector<float> U((size_t)e->m * f), V((size_t)e->n * f), vals(e->nnz);
    int rc = als_get_factors(e, 0, U.data(), e->m, f);
    if (!rc) rc = als_get_factors(e, 1, V.data(), e->n, f);
    if (rc) return rc;
    if (e->nnz > 0) {
        HIPCHK(e, hipMemcpyAsync(vals.data(), e->side[0].val, e->nnz * sizeof(float), hipMemcpyDeviceToHost, e->stream));
        HIPCHK(e, hipStreamSynchronize(e->stream));
    }
    double s = 0.0;
    for (float v : vals) s += std::fabs((double)v);
    float mean = e->nnz > 0 ? (float)(s / (double)e->nnz) : 0.f;
    std::ofstream out(path, std::ios::binary);
    if (!out) return set_err(e, ALS_E_IO, std::string(path) + ": cannot open");
    int hdr[4] = {0, (int)e->m, (int)e->n, f};
    out.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
    out.write(reinterpret_cast<const char *>(&mean), sizeof(mean));
    out.write(reinterpret_cast<const char *>(U.data()), U.size() * sizeof(float));
    out.write(reinterpret_cast<const char *>(V.data()), V.size() * sizeof(float));
    if (!out) return set_err(e, ALS_E_IO, std::string(path) + ": write failed");
    return ALS_OK;
}

}  // extern "C"

namespace {

// Bytes of the saved factorisation of one side: NT tiles of 256 floats per row.
int64_t factor_save_bytes(const als_engine *e, int sd) {
    const int64_t rows = sd == 0 ? e->m : e->n;
    return std::max<int64_t>(1, rows) * (int64_t)(e->T * (e->T + 1) / 2) * 256 * (int64_t)sizeof(float);
}

int forward_impl(als_engine *e, int sd, const float *d_Y, float *d_X, const float *d_lambda_rows, float *Ls) {
    if (!e->have_train) return set_err(e, ALS_E_STATE, "set the train data first");
    if (sd != 0 && sd != 1) return set_err(e, ALS_E_INVALID, "bad side");
    if (e->solve == SOLVE_CG) return set_err(e, ALS_E_UNSUPPORTED, "the adjoint needs the Cholesky solver");
    const int64_t cols = sd == 0 ? e->n : e->m;
    int rc = ALS_OK;
    const uint16_t *Ys = nullptr;
    if (e->prec != PREC_FP32) {
        size_t need = (size_t)std::max<int64_t>(1, cols * e->ld) * 2;
        if (e->tmp16_bytes < need) {
            dev_free(e, e->tmp16);
            if ((rc = dev_alloc(e, &e->tmp16, need / 2))) return rc;
            e->tmp16_bytes = need;
        }
        HIPCHK(e, launch_to16(e->prec, d_Y, e->tmp16, cols * e->ld, e->stream));
        Ys = e->tmp16;
    }
    return launch_side(e, sd, d_Y, Ys, d_X, nullptr, d_lambda_rows, Ls, e->stream);
}

int adjoint_impl(als_engine *e, int sd, const float *d_Y, const float *d_X, const float *Ls, const float *d_gbar,
                 float *d_xbar, float *d_lambda_bar, float *d_ybar) {
    const int64_t rows = sd == 0 ? e->m : e->n, cols = sd == 0 ? e->n : e->m;
    AdjointArgs a{};
    a.ptr = e->side[sd].ptr;
    a.X = d_X;
    a.Lsave = Ls;
    a.Gbar = d_gbar;
    a.Xbar = d_xbar;
    a.lambda_bar = d_lambda_bar;
    a.rows = (int)rows;
    a.f = e->cfg.f;
    if (rows > 0) HIPCHK(e, launch_adjoint_solve(e->T, a, e->stream));
    if (d_ybar && cols > 0) {
        if (!e->yb_parts[sd]) {
            // parts of at most YBAR_PART ratings of every column, columns in order
            const std::vector<int> &hp = e->hptr[1 - sd];
            std::vector<int4> parts;
            std::vector<int> colpart(cols + 1);
            for (int64_t c = 0; c < cols; ++c) {
                colpart[c] = (int)parts.size();
                for (int64_t k = hp[c]; k < hp[c + 1]; k += YBAR_PART)
                    parts.push_back(int4{(int)c, (int)k, (int)std::min<int64_t>(hp[c + 1], k + YBAR_PART), 0});
            }
            colpart[cols] = (int)parts.size();
            // kept only once every buffer is allocated and filled
            int4 *dparts = nullptr;
            int *dcolpart = nullptr;
            float *dpartial = nullptr;
            int rc = dev_alloc(e, &dparts, parts.size());
            if (!rc) rc = dev_alloc(e, &dcolpart, colpart.size());
            if (!rc) rc = dev_alloc(e, &dpartial, std::max<size_t>(1, parts.size()) * e->ld);
            hipError_t st = hipSuccess;
            if (!rc && !parts.empty())
                st = hipMemcpyAsync(dparts, parts.data(), parts.size() * sizeof(int4), hipMemcpyHostToDevice,
                                    e->stream);
            if (!rc && st == hipSuccess)
                st = hipMemcpyAsync(dcolpart, colpart.data(), colpart.size() * sizeof(int), hipMemcpyHostToDevice,
                                    e->stream);
            if (!rc && st == hipSuccess) st = hipStreamSynchronize(e->stream);
            if (rc || st != hipSuccess) {
                dev_free(e, dparts);
                dev_free(e, dcolpart);
                dev_free(e, dpartial);
                return rc ? rc : set_err(e, ALS_E_HIP, std::string("adjoint work table: ") + hipGetErrorString(st));
            }
            e->yb_parts[sd] = dparts;
            e->yb_colpart[sd] = dcolpart;
            e->yb_partial[sd] = dpartial;
            e->yb_nparts[sd] = (int)parts.size();
        }
        YbarArgs y{};
        y.parts = e->yb_parts[sd];
        y.colpart = e->yb_colpart[sd];
        y.partial = e->yb_partial[sd];
        y.nparts = e->yb_nparts[sd];
        const Side &other = e->side[1 - sd];  // CSC of this side's matrix = CSR of the other side
        y.cptr = other.ptr;
        y.cidx = other.idx;
        y.cval = other.val;
        y.X = d_X;
        y.Xbar = d_xbar;
        y.Y = d_Y;
        y.Ybar = d_ybar;
        y.ncols = (int)cols;
        y.f = e->cfg.f;
        HIPCHK(e, launch_ybar(e->T, y, e->stream));
    }
    return ALS_OK;
}

}  // namespace

extern "C" {

int als_forward_dev(als_engine *e, int sd, const float *d_Y, float *d_X, const float *d_lambda_rows, int save) {
    if (!e || !d_Y || !d_X) return ALS_E_INVALID;
    if (sd != 0 && sd != 1) return set_err(e, ALS_E_INVALID, "bad side");
    float *Ls = nullptr;
    if (save) {
        if (!e->have_train) return set_err(e, ALS_E_STATE, "set the train data first");
        const size_t need = (size_t)factor_save_bytes(e, sd);
        if (e->Lsave_bytes < need) {
            dev_free(e, e->Lsave);
            e->Lsave = nullptr;
            e->Lsave_bytes = 0;
            int rc = dev_alloc(e, &e->Lsave, need / sizeof(float));
            if (rc) return rc;
            e->Lsave_bytes = need;
        }
        Ls = e->Lsave;
    }
    const int rc = forward_impl(e, sd, d_Y, d_X, d_lambda_rows, Ls);
    // the engine keeps one saved factorisation: the last saving call's side
    if (save) e->Lsave_side = rc == ALS_OK ? sd : -1;
    return rc;
}

int als_adjoint_dev(als_engine *e, int sd, const float *d_Y, const float *d_X, const float *d_gbar, float *d_xbar,
                    float *d_lambda_bar, float *d_ybar) {
    if (!e || !d_Y || !d_X || !d_gbar || !d_xbar || !d_lambda_bar) return ALS_E_INVALID;
    if (sd != 0 && sd != 1) return set_err(e, ALS_E_INVALID, "bad side");
    if (!e->Lsave || e->Lsave_side != sd)
        return set_err(e, ALS_E_STATE, "the engine's saved factorisation is not of this side: run als_forward_dev "
                                       "with save_factor on this side first (or keep one per call with "
                                       "als_forward_dev_ws / als_adjoint_dev_ws)");
    return adjoint_impl(e, sd, d_Y, d_X, e->Lsave, d_gbar, d_xbar, d_lambda_bar, d_ybar);
}

int als_factor_save_bytes(als_engine *e, int sd, int64_t *bytes) {
    if (!e || !bytes) return ALS_E_INVALID;
    if (!e->have_train) return set_err(e, ALS_E_STATE, "set the train data first");
    if (sd != 0 && sd != 1) return set_err(e, ALS_E_INVALID, "bad side");
    *bytes = factor_save_bytes(e, sd);
    return ALS_OK;
}

int als_forward_dev_ws(als_engine *e, int sd, const float *d_Y, float *d_X, const float *d_lambda_rows,
                       float *d_save, int64_t save_bytes) {
    if (!e || !d_Y || !d_X || !d_save) return ALS_E_INVALID;
    if (!e->have_train) return set_err(e, ALS_E_STATE, "set the train data first");
    if (sd != 0 && sd != 1) return set_err(e, ALS_E_INVALID, "bad side");
    if (save_bytes < factor_save_bytes(e, sd))
        return set_err(e, ALS_E_INVALID, "save buffer smaller than als_factor_save_bytes()");
    return forward_impl(e, sd, d_Y, d_X, d_lambda_rows, d_save);
}

int als_adjoint_dev_ws(als_engine *e, int sd, const float *d_Y, const float *d_X, const float *d_save,
                       int64_t save_bytes, const float *d_gbar, float *d_xbar, float *d_lambda_bar, float *d_ybar) {
    if (!e || !d_Y || !d_X || !d_save || !d_gbar || !d_xbar || !d_lambda_bar) return ALS_E_INVALID;
    if (!e->have_train) return set_err(e, ALS_E_STATE, "set the train data first");
    if (sd != 0 && sd != 1) return set_err(e, ALS_E_INVALID, "bad side");
    if (save_bytes < factor_save_bytes(e, sd))
        return set_err(e, ALS_E_INVALID, "save buffer smaller than als_factor_save_bytes()");
    return adjoint_impl(e, sd, d_Y, d_X, d_save, d_gbar, d_xbar, d_lambda_bar, d_ybar);
}

}  // extern "C"

// Applies a schedule directly (a result of als_autotune, say); values <= 0
// keep the current ones.
extern "C" int als_set_schedule(als_engine *e, int sd, int64_t split_len, int64_t batch_work, int restage_max) {
    if (!e) return ALS_E_INVALID;
    if (!e->have_train) return set_err(e, ALS_E_STATE, "set the train data first");
    if (sd != 0 && sd != 1) return set_err(e, ALS_E_INVALID, "bad side");
    Side &s = e->side[sd];
    drop_graph(e);
    if (restage_max >= 0) s.restage_max = restage_max;
    if (split_len > 0 || batch_work > 0)
        return build_side(e, s, s.rows, e->hptr[sd].data(), split_len > 0 ? split_len : s.split_len,
                          batch_work > 0 ? batch_work : s.batch_work);
    return ALS_OK;
}

// Selects the update kernel variant of one side (0: default, 1: one wave per
// SIMD where it is built for the engine's solver and tile count).
extern "C" int als_set_kernel_variant(als_engine *e, int sd, int variant) {
    if (!e) return ALS_E_INVALID;
    if (sd != 0 && sd != 1) return set_err(e, ALS_E_INVALID, "bad side");
    if (variant != 0 && variant != 1) return set_err(e, ALS_E_INVALID, "variant must be 0 or 1");
    if (variant == 1 && e->grid_v1 == 0)
        return set_err(e, ALS_E_UNSUPPORTED, "no one-wave kernel for this solver and tile count");
    drop_graph(e);
    e->side[sd].variant = variant;
    return ALS_OK;
}

// Persistent waves of the default update kernel on one side (1 ..
// grid_max); the schedule is rebuilt for that many waves.
extern "C" int als_set_grid(als_engine *e, int sd, int grid) {
    if (!e) return ALS_E_INVALID;
    if (!e->have_train) return set_err(e, ALS_E_STATE, "set the train data first");
    if (sd != 0 && sd != 1) return set_err(e, ALS_E_INVALID, "bad side");
    if (grid < 1 || grid > e->grid_max)
        return set_err(e, ALS_E_INVALID, "grid must be between 1 and " + std::to_string(e->grid_max));
    drop_graph(e);
    Side &s = e->side[sd];
    s.grid = grid;
    return build_side(e, s, s.rows, e->hptr[sd].data(), s.split_len, s.batch_work);
}

// Rows of one side with fewer than `limit` ratings (and empty rows) take the
// dual system; 0 sends every row through the f x f system. At most the
// engine's dual_max.
extern "C" int als_set_dual_limit(als_engine *e, int sd, int limit) {
    if (!e) return ALS_E_INVALID;
    if (!e->have_train) return set_err(e, ALS_E_STATE, "set the train data first");
    if (sd != 0 && sd != 1) return set_err(e, ALS_E_INVALID, "bad side");
    if (limit < 0 || limit > e->dual_max)
        return set_err(e, ALS_E_INVALID, "dual limit must be between 0 and " + std::to_string(e->dual_max));
    drop_graph(e);
    Side &s = e->side[sd];
    s.dual_limit = limit;
    return build_side(e, s, s.rows, e->hptr[sd].data(), s.split_len, s.batch_work);
}

// How one side's 16-bit Gram gathers its rows: 1 straight into LDS (LDS-DMA,
// two chunks in flight; built for gram_dma_built tile counts), 0 through
// registers (one chunk in flight). The results are identical.
extern "C" int als_set_gram_gather(als_engine *e, int sd, int mode) {
    if (!e) return ALS_E_INVALID;
    if (sd != 0 && sd != 1) return set_err(e, ALS_E_INVALID, "bad side");
    if (mode != 0 && mode != 1) return set_err(e, ALS_E_INVALID, "mode must be 0 or 1");
    if (mode == 1 && !gram_dma_built(e->T, e->prec))
        return set_err(e, ALS_E_UNSUPPORTED, "no LDS-DMA Gram for this precision and tile count");
    drop_graph(e);
    e->side[sd].gram_dma = mode;
    return ALS_OK;
}

// Autotuner: times one half-sweep per candidate schedule (split length x
// batch size, around the cost-model defaults, under each kernel variant the
// engine has) on the engine's own data and keeps the fastest; then the
// update kernel's grid (fewer persistent waves
// than the occupancy limit can gather faster: the items half of the
// netflix-shaped workload runs 4% faster at six waves per CU than at
// eight), the shape of the light rows' systems (the rating count below which
// a row solves its n x n dual system in ceil((n+1)/16) tiles instead of the
// f x f one in T tiles), the restage threshold and the kernel variant. The side's factors and the failure count are restored
// after every trial, so tuning leaves the model state unchanged.
extern "C" int als_autotune(als_engine *e, int sides, als_tune_result *out) {
    if (!e) return ALS_E_INVALID;
    if (!e->have_train) return set_err(e, ALS_E_STATE, "set the train data first");
    if (sides & ~3) return set_err(e, ALS_E_INVALID, "sides: bit 0 users, bit 1 items");
    als_tune_result res{};
    drop_graph(e);
    int rc = ALS_OK;
    hipEvent_t ev[2] = {nullptr, nullptr};
    HIPCHK(e, hipEventCreate(&ev[0]));
    HIPCHK(e, hipEventCreate(&ev[1]));
    for (int sd = 0; sd < 2 && !rc; sd++) {
        Side &s = e->side[sd];
        if (!(sides & (1 << sd)) || s.rows == 0) {
            res.split_len[sd] = s.split_len;
            res.batch_work[sd] = s.batch_work;
            res.restage_max[sd] = s.restage_max;
            res.kernel_variant[sd] = s.variant;
            res.grid[sd] = s.grid;
            res.dual_limit[sd] = s.dual_limit;
            res.gram_gather[sd] = gram_dma_built(e->T, e->prec) ? s.gram_dma : 0;
            continue;
        }
        const int64_t rows = s.rows;
        const int *hp = e->hptr[sd].data();
        int64_t maxrow = 1;
        for (int64_t r = 0; r < rows; r++) maxrow = std::max<int64_t>(maxrow, hp[r + 1] - hp[r]);
        const int64_t sl0 = s.split_len, bw0 = s.batch_work;
        std::vector<std::pair<int64_t, int64_t>> cand;
        for (int64_t sl : {sl0, sl0 / 2, sl0 * 2, sl0 * 4, maxrow + 1})
            for (int64_t bw : {bw0, bw0 / 4, bw0 * 4}) {
                const auto c = std::make_pair(std::max<int64_t>(256, std::min<int64_t>(sl, INT32_MAX / 2)),
                                              std::max<int64_t>(1, std::min<int64_t>(bw, INT32_MAX / 2)));
                if (std::find(cand.begin(), cand.end(), c) == cand.end()) cand.push_back(c);
            }
        // save what a trial overwrites
        float *X = sd == 0 ? e->U : e->V;
        uint16_t *Xs = sd == 0 ? e->Us : e->Vs;
        const size_t cnt = (size_t)std::max<int64_t>(1, rows * e->ld);
        float *xsave = nullptr;
        uint16_t *xssave = nullptr;
        int *fsave = nullptr;
        if ((rc = dev_alloc(e, &xsave, cnt)) || (Xs && (rc = dev_alloc(e, &xssave, cnt))) ||
            (rc = dev_alloc(e, &fsave, 1)))
            break;
        HIPCHK(e, hipMemcpyAsync(xsave, X, cnt * sizeof(float), hipMemcpyDeviceToDevice, e->stream));
        if (Xs) HIPCHK(e, hipMemcpyAsync(xssave, Xs, cnt * 2, hipMemcpyDeviceToDevice, e->stream));
        HIPCHK(e, hipMemcpyAsync(fsave, e->fail, sizeof(int), hipMemcpyDeviceToDevice, e->stream));
        auto restore = [&]() -> int {
            HIPCHK(e, hipMemcpyAsync(X, xsave, cnt * sizeof(float), hipMemcpyDeviceToDevice, e->stream));
            if (Xs) HIPCHK(e, hipMemcpyAsync(Xs, xssave, cnt * 2, hipMemcpyDeviceToDevice, e->stream));
            HIPCHK(e, hipMemcpyAsync(e->fail, fsave, sizeof(int), hipMemcpyDeviceToDevice, e->stream));
            return ALS_OK;
        };
        // one candidate: a run that warms the caches, then the faster of two timed runs
        auto trial = [&](float &t) -> int {
            t = 1e30f;
            for (int rep = 0; rep < 3; rep++) {
                int r;
                if ((r = restore())) return r;
                HIPCHK(e, hipEventRecord(ev[0], e->stream));
                if ((r = update_side(e, sd))) return r;
                HIPCHK(e, hipEventRecord(ev[1], e->stream));
                HIPCHK(e, hipEventSynchronize(ev[1]));
                float ms = 0.f;
                HIPCHK(e, hipEventElapsedTime(&ms, ev[0], ev[1]));
                if (rep > 0) t = std::min(t, ms);
            }
            return ALS_OK;
        };
        // split length x batch size, under each kernel variant the engine
        // has (the best split length depends on it: a variant's waves take
        // a split row's parts and finishes at their own rate)
        float best = 1e30f;
        size_t besti = 0;
        const int sv0 = s.variant;
        int bestsv = sv0;
        for (int v : {sv0, 1 - sv0}) {
            if (v != sv0 && e->grid_v1 == 0) continue;
            s.variant = v;
            for (size_t i = 0; i < cand.size() && !rc; i++) {
                if ((rc = build_side(e, s, rows, hp, cand[i].first, cand[i].second))) break;
                float t;
                if ((rc = trial(t))) break;
                if (i == 0 && v == sv0) res.default_ms[sd] = t;
                if (t < best) {
                    best = t;
                    besti = i;
                    bestsv = v;
                }
            }
            if (rc) break;
        }
        s.variant = bestsv;
        if (!rc) rc = build_side(e, s, rows, hp, cand[besti].first, cand[besti].second);
        // then the grid: the occupancy limit and up to half as many waves per
        // CU, the schedule rebuilt for each
        if (!rc) {
            const int pc = e->grid_max / e->num_cu, g0 = s.grid;
            int bestg = g0;
            for (int k = 1; k <= 4 && !rc; k++) {
                const int g = (pc - k) * e->num_cu;
                if (2 * (pc - k) < pc || pc - k < 1) break;
                s.grid = g;
                float t;
                if ((rc = build_side(e, s, rows, hp, s.split_len, s.batch_work)) || (rc = trial(t))) break;
                if (t < best) {
                    best = t;
                    bestg = g;
                }
            }
            s.grid = bestg;
            if (!rc) rc = build_side(e, s, rows, hp, s.split_len, s.batch_work);
        }
        // then the dual limit: the built maximum and one or two tiles fewer
        if (!rc && e->dual_max > 0) {
            const int d0 = s.dual_limit;
            int bestd = d0;
            for (int dl : {e->dual_max, e->dual_max - 16, e->dual_max - 32}) {
                if (dl < 16 || dl == d0) continue;
                s.dual_limit = dl;
                float t;
                if ((rc = build_side(e, s, rows, hp, s.split_len, s.batch_work)) || (rc = trial(t))) break;
                if (t < best) {
                    best = t;
                    bestd = dl;
                }
            }
            s.dual_limit = bestd;
            if (!rc) rc = build_side(e, s, rows, hp, s.split_len, s.batch_work);
        }
        // then the refinement's restage threshold on the chosen schedule
        if (!rc && e->solve == SOLVE_CHOL_REFINE && e->cfg.refine_steps > 0) {
            const int r0 = s.restage_max;
            int bestr = r0;
            float bestt = 1e30f;
            for (int rm : {r0, 0, 32, 64, 128, 512, 1024, 1 << 30}) {
                s.restage_max = rm;
                float t;
                if ((rc = trial(t))) break;
                if (t < bestt) {
                    bestt = t;
                    bestr = rm;
                }
            }
            s.restage_max = bestr;
            best = std::min(best, bestt);
        }
        // then the kernel variant (one wave per SIMD against two) and the Gram's
        // gather (straight to LDS or through registers), together, on the
        // chosen schedule
        const bool dma_built = gram_dma_built(e->T, e->prec);
        if (!rc && (e->grid_v1 > 0 || dma_built)) {
            const int v0 = s.variant, d0 = s.gram_dma;
            int bestv = v0, bestd = d0;
            float bestvt = 1e30f;
            for (int v : {v0, 1 - v0}) {
                if (v != v0 && e->grid_v1 == 0) continue;
                for (int dm : {d0, 1 - d0}) {
                    if (dm != d0 && !dma_built) continue;
                    s.variant = v;
                    s.gram_dma = dm;
                    float t;
                    if ((rc = trial(t))) break;
                    if (t < bestvt) {
                        bestvt = t;
                        bestv = v;
                        bestd = dm;
                    }
                }
                if (rc) break;
            }
            s.variant = bestv;
            s.gram_dma = bestd;
            best = std::min(best, bestvt);
        }
        if (!rc) rc = restore();
        if (!rc) HIPCHK(e, hipStreamSynchronize(e->stream));
        res.restage_max[sd] = s.restage_max;
        res.kernel_variant[sd] = s.variant;
        res.grid[sd] = s.grid;
        res.dual_limit[sd] = s.dual_limit;
        res.gram_gather[sd] = dma_built ? s.gram_dma : 0;
        dev_free(e, xsave);
        dev_free(e, xssave);
        dev_free(e, fsave);
        res.best_ms[sd] = best;
        res.candidates[sd] = (int)cand.size();
        res.split_len[sd] = s.split_len;
        res.batch_work[sd] = s.batch_work;
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (out) *out = res;
    return rc;
}

// ------------------------------------------------------------------ groups
// Several engines (models of different f) over one rating matrix, updated
// per half-sweep by ONE graph launch: each model's own kernels (update and
// dual), forked onto a stream per model so that they run concurrently and
// fill each other's tails. One kernel for all models was tried first: its
// register allocation is set by the largest f, and for f = 32/64/128 it
// took 29.8 ms per sweep against 26.4 ms for three separate engines.
struct als_group {
    std::vector<als_engine *> eng;
    std::vector<hipStream_t> streams;  // one per model
    std::vector<hipEvent_t> pre;       // each engine's stream, before the launch
    std::vector<hipEvent_t> join;      // each model's branch of the graph
    hipEvent_t fork = nullptr, post = nullptr;
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    std::vector<char> sig[2];          // launch arguments the cached graph was captured with
    std::string err;
};

namespace {
int gset_err(als_group *g, int code, const std::string &msg) {
    g->err = msg;
    return code;
}
void group_free(als_group *g) {
    if (!g) return;
    if (!g->eng.empty()) (void)hipSetDevice(g->eng[0]->cfg.device);
    for (hipGraphExec_t &x : g->exec)
        if (x) (void)hipGraphExecDestroy(x);
    for (hipStream_t st : g->streams) (void)hipStreamSynchronize(st);
    for (hipStream_t st : g->streams) (void)hipStreamDestroy(st);
    for (hipEvent_t ev : g->pre) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : g->join) (void)hipEventDestroy(ev);
    if (g->fork) (void)hipEventDestroy(g->fork);
    if (g->post) (void)hipEventDestroy(g->post);
    delete g;
}

// Everything a half-sweep launch of the group depends on, as bytes: the
// cached graph is replayed only while this stays the same (schedules,
// buffers and grids are baked into its kernel nodes).
std::vector<char> group_signature(als_group *g, int sd) {
    std::vector<char> sig;
    auto put = [&](const void *p, size_t n) { sig.insert(sig.end(), (const char *)p, (const char *)p + n); };
    for (als_engine *e : g->eng) {
        const Side &s = e->side[sd];
        const UpdateArgs a = sd == ALS_SIDE_USERS
                                 ? side_args(e, s, e->V, e->Vs, e->U, e->Us, e->side[0].lambda_rows, nullptr)
                                 : side_args(e, s, e->U, e->Us, e->V, e->Vs, e->side[1].lambda_rows, nullptr);
        put(&a, sizeof(a));
        const int v[] = {e->T, e->prec, e->solve, s.grid, s.variant, e->grid_v1, e->lds, s.nbatches, s.dnbatches,
                         e->dual_grid, e->dual_lds};
        put(v, sizeof(v));
        put(&s.dunits, sizeof(s.dunits));
        put(&s.dbatch, sizeof(s.dbatch));
    }
    return sig;
}
}  // namespace

extern "C" int als_group_create(als_engine *const *engines, int count, als_group **out) {
    if (!engines || !out || count < 1 || count > MAX_MODELS) return ALS_E_INVALID;
    *out = nullptr;
    als_engine *e0 = engines[0];
    for (int i = 0; i < count; i++) {
        const als_engine *e = engines[i];
        if (!e || !e->have_train) return ALS_E_STATE;
        if (e->m != e0->m || e->n != e0->n || e->nnz != e0->nnz) return ALS_E_INVALID;
        if (e->cfg.device != e0->cfg.device) return ALS_E_INVALID;
        for (int j = 0; j < i; j++)
            if (engines[j] == e) return ALS_E_INVALID;
    }
    auto *g = new als_group;
    g->eng.assign(engines, engines + count);
    // the heaviest model first: its branch is forked first
    std::stable_sort(g->eng.begin(), g->eng.end(), [](const als_engine *a, const als_engine *b) { return a->T > b->T; });
    if (hipSetDevice(e0->cfg.device) != hipSuccess) { group_free(g); return ALS_E_HIP; }
    bool ok = hipEventCreateWithFlags(&g->fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&g->post, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; ok && i < count; i++) {
        hipStream_t st = nullptr;
        hipEvent_t a = nullptr, b = nullptr;
        ok = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
        if (ok) g->streams.push_back(st);
        ok = ok && hipEventCreateWithFlags(&a, hipEventDisableTiming) == hipSuccess;
        if (ok) g->pre.push_back(a);
        ok = ok && hipEventCreateWithFlags(&b, hipEventDisableTiming) == hipSuccess;
        if (ok) g->join.push_back(b);
    }
    if (!ok) { group_free(g); return ALS_E_HIP; }
    *out = g;
    return ALS_OK;
}

extern "C" void als_group_destroy(als_group *g) { group_free(g); }

extern "C" const char *als_group_last_error(const als_group *g) { return g ? g->err.c_str() : "null group"; }

extern "C" int als_group_update(als_group *g, int sd) {
    if (!g) return ALS_E_INVALID;
    if (sd != ALS_SIDE_USERS && sd != ALS_SIDE_ITEMS) return gset_err(g, ALS_E_INVALID, "side must be users or items");
    als_engine *e0 = g->eng[0];
    hipStream_t s0 = e0->stream;
    const int nm = (int)g->eng.size();
    for (als_engine *e : g->eng) drop_graph(e);
#define GCHK(expr)                                                                                  \
    do {                                                                                            \
        const hipError_t _st = (expr);                                                              \
        if (_st != hipSuccess) return gset_err(g, ALS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_st)); \
    } while (0)
    std::vector<char> sig = group_signature(g, sd);
    if (!g->exec[sd] || sig != g->sig[sd]) {
        if (g->exec[sd]) {
            (void)hipGraphExecDestroy(g->exec[sd]);
            g->exec[sd] = nullptr;
        }
        // capture: fork from s0 to one stream per model, each model's own
        // launches on its stream, join back into s0
        GCHK(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
        int rc = ALS_OK;
        const als_engine *failed = nullptr;
        hipError_t st = hipEventRecord(g->fork, s0);
        for (int k = 0; st == hipSuccess && rc == ALS_OK && k < nm; k++) {
            st = hipStreamWaitEvent(g->streams[k], g->fork, 0);
            if (st == hipSuccess) rc = launch_own_side(g->eng[k], sd, g->streams[k], false);
            if (rc != ALS_OK) failed = g->eng[k];
            if (st == hipSuccess && rc == ALS_OK) st = hipEventRecord(g->join[k], g->streams[k]);
            if (st == hipSuccess && rc == ALS_OK) st = hipStreamWaitEvent(s0, g->join[k], 0);
        }
        hipGraph_t graph = nullptr;
        const hipError_t se = hipStreamEndCapture(s0, &graph);
        if (rc != ALS_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return gset_err(g, rc, "group capture: " + failed->err);
        }
        if (st != hipSuccess || se != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return gset_err(g, ALS_E_HIP, std::string("group capture: ") + hipGetErrorString(st != hipSuccess ? st : se));
        }
        const hipError_t ie = hipGraphInstantiate(&g->exec[sd], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) {
            g->exec[sd] = nullptr;
            return gset_err(g, ALS_E_HIP, std::string("group graph: ") + hipGetErrorString(ie));
        }
        g->sig[sd] = std::move(sig);
    }
    // order after every engine's pending work, and every engine's later
    // work after the group
    for (int k = 0; k < nm; k++) {
        if (g->eng[k]->stream == s0) continue;
        GCHK(hipEventRecord(g->pre[k], g->eng[k]->stream));
        GCHK(hipStreamWaitEvent(s0, g->pre[k], 0));
    }
    GCHK(hipGraphLaunch(g->exec[sd], s0));
    GCHK(hipEventRecord(g->post, s0));
    for (int k = 0; k < nm; k++)
        if (g->eng[k]->stream != s0) GCHK(hipStreamWaitEvent(g->eng[k]->stream, g->post, 0));
#undef GCHK
    return ALS_OK;
}

extern "C" int als_group_sweep(als_group *g) {
    int rc = als_group_update(g, ALS_SIDE_USERS);
    return rc ? rc : als_group_update(g, ALS_SIDE_ITEMS);
}

// Batched CG on caller device buffers, synchronous like the reference's
// updateXWithCGHost (cumf_als_cg/cg.h:30-32, cg.cu:683-688): the kernel is
// enqueued on hip_stream (null: the default stream) of the current device,
// and the call returns once that stream has finished it.
static int cg_update_x(const float *A, const uint16_t *A16, float *x, const float *b, int batch, int f, int cg_iters,
                       float tol, void *hip_stream, bool round_a16 = false) {
    if (batch < 0 || f < 1 || f > 143 || cg_iters < 0 || !(tol >= 0.f)) return ALS_E_INVALID;
    if (batch == 0) return ALS_OK;
    if ((!A && !A16) || !x || !b) return ALS_E_INVALID;
    CgBatchArgs a{};
    a.A = A;
    a.A16 = A16;
    a.round_a16 = round_a16 ? 1 : 0;
    a.vec = f % 4 == 0 && (A ? reinterpret_cast<uintptr_t>(A) % 16 == 0 : reinterpret_cast<uintptr_t>(A16) % 8 == 0);
    a.x = x;
    a.b = b;
    a.batch = batch;
    a.f = f;
    a.iters = cg_iters;
    a.tol = tol;
    hipStream_t s = reinterpret_cast<hipStream_t>(hip_stream);
    if (launch_cg_batch((f + 16) / 16, a, s) != hipSuccess) return ALS_E_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return ALS_E_HIP;
    return ALS_OK;
}

extern "C" int als_cg_update_x(const float *A, float *x, const float *b, int batch, int f, int cg_iters, float tol,
                               void *hip_stream) {
    return cg_update_x(A, nullptr, x, b, batch, f, cg_iters, tol, hip_stream);
}

extern "C" int als_cg_update_x_fp16(const uint16_t *A, float *x, const float *b, int batch, int f, int cg_iters,
                                    float tol, void *hip_stream) {
    return cg_update_x(nullptr, A, x, b, batch, f, cg_iters, tol, hip_stream);
}

extern "C" int als_cg_update_x_tt_fp16(const float *A, float *x, const float *b, int batch, int f, int cg_iters,
                                       float tol, void *hip_stream) {
    return cg_update_x(A, nullptr, x, b, batch, f, cg_iters, tol, hip_stream, true);
}

extern "C" int als_schedule_preview(int64_t rows, const int *ptr, int split_len, int batch_work, int grid,
                                    int gram_prec, als_schedule_stats *out, int *units_out, int64_t max_units) {
    if (!ptr || !out || rows < 0 || grid <= 0 || gram_prec < ALS_PREC_FP32 || gram_prec > ALS_PREC_FP16)
        return ALS_E_INVALID;
    for (int64_t r = 0; r < rows; r++)
        if (ptr[r + 1] < ptr[r]) return ALS_E_INVALID;
    Schedule sc = build_schedule(rows, ptr, split_len, batch_work, grid, gram_prec);
    out->nunits = (int64_t)sc.units.size();
    out->nsplit_rows = (int64_t)sc.splits.size();
    out->nbatches = (int64_t)sc.batch_begin.size() - 1;
    out->nslots = sc.nslots;
    out->split_len = sc.split_len;
    out->batch_work = sc.batch_work;
    if (units_out)
        for (int64_t u = 0; u < (int64_t)sc.units.size() && u < max_units; u++) {
            units_out[4 * u + 0] = sc.units[u].row;
            units_out[4 * u + 1] = sc.units[u].kb;
            units_out[4 * u + 2] = sc.units[u].ke;
            units_out[4 * u + 3] = sc.units[u].split;
        }
    return ALS_OK;
}
