// Kernel argument blocks and launchers shared by the engine (host) and the
// device code. Internal to the library: the public boundary is include/als.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace als {

enum Prec { PREC_FP32 = 0, PREC_BF16 = 1, PREC_FP16 = 2 };
// SOLVE_CHOL_WARM: one warm-started refinement step and no other (its own
// kernel: without the classic residual code the registers fit without spills).
enum SolveKind { SOLVE_CHOL = 0, SOLVE_CG = 1, SOLVE_CHOL_REFINE = 2, SOLVE_CHOL_WARM = 3 };
constexpr int MAX_T = 9;  // augmented dimension f + 1 <= 16 * MAX_T

struct UpdateArgs {
    const int *ptr;             // CSR of the side being updated (int32, reference layout)
    const int *idx;
    const float *val;
    const float *Y;             // fixed factors, FP32, row stride 16*T
    const uint16_t *Ys;         // fixed factors, 16-bit shadow, row stride 16*T (16-bit Gram only)
    float *X;                   // updated factors, FP32, row stride 16*T (CG: warm start in)
    uint16_t *Xs;               // 16-bit shadow of X to refresh (nullable)
    const float *lambda_rows;   // per-row lambda (nullable -> lambda)
    float lambda;
    int f;
    const int4 *units;          // {row, kb, ke, split} work units, heaviest first
    const int4 *splits;         // {slot_base, nparts, part_len, -} per split row
    int *counters;              // arrivals per split row (0 between launches)
    float *partial;             // partial Gram tiles, NT*256 floats per slot
    const int *batch_begin;     // batch b = units[batch_begin[b] .. batch_begin[b+1])
    int nbatches;
    int *queue;                 // [0] next batch, [1] waves that left
    int cg_iters;
    float cg_tol;
    int cg_fp16_a;              // CG on an FP16-rounded A
    int refine;                 // FP64-residual refinement steps (Cholesky)
    int refine_warm;            // first refinement step from the row's previous value in X
    int restage_max;            // whole rows up to this many ratings form the residual by
                                // restaging their ratings; longer ones use the tile copy
    int *fail_count;            // rows whose solve broke down (x set to 0)
    float *Lsave;               // Cholesky factor tiles per row, NT*256 floats (nullable)
    float *scratch;             // per-wave copy of A for refinement, NT*256 floats per workgroup
    const float *zero32;        // one row of zeros (16*T floats) read by unused chunk slots
    const uint16_t *zero16;     // same, 16-bit
    int ys_bytes;               // bytes of the 16-bit Y shadow when below 2^31 (the pipelined
                                // Gram then gathers with 32-bit buffer offsets), else 0
    int gram_dma;               // 1: the 16-bit Gram gathers straight to LDS where built
                                // (gram_dma_built), 0: through registers
};

// Tile counts and precisions whose update kernels carry the LDS-DMA Gram:
// a 16-bit Gram, at most 8 tiles (a chunk's gather tasks fit a wave) and
// T % 4 != 0 (rows of 16T halves keep a transposing read on distinct banks).
constexpr bool gram_dma_built(int T, int prec) { return prec != PREC_FP32 && T <= 8 && T % 4 != 0; }

// Several models (different f) on one rating matrix, one launch: batch b of
// the combined queue is batch b - first[k] of model k (first[k] <= b < first[k+1]).
constexpr int MAX_MODELS = 4;  // engines per group (als_group_*)

struct AdjointArgs {
    const int *ptr;             // CSR of the side that was updated
    const float *X;             // forward solution, stride 16*T
    const float *Lsave;         // packed factors from the forward pass
    const float *Gbar;          // dL/dX, stride 16*T
    float *Xbar;                // A^-1 Gbar, stride 16*T
    float *lambda_bar;          // -n_u Xbar_u . X_u
    int rows;
    int f;
};

// Batched CG on caller-provided dense systems (the reference's
// updateXWithCGHost boundary): A batch x f x f (symmetric, either order),
// x and b batch x f, all densely packed.
struct CgBatchArgs {
    const float *A;             // FP32 systems, or nullptr when A16 is set
    const uint16_t *A16;        // FP16 systems
    int round_a16;              // FP32 systems whose entries are rounded to FP16 on load
    int vec;                    // f % 4 == 0 and A 16-byte (A16 8-byte) aligned: a lane
                                // loads its four entries of a tile in one load
    float *x;                   // warm start in, solution out
    const float *b;
    int batch;
    int f;
    int iters;
    float tol;
};

struct YbarArgs {
    const int *cptr;            // CSC of the same matrix (columns = fixed-factor rows)
    const int *cidx;
    const float *cval;
    const float *X, *Xbar, *Y;  // stride 16*T
    float *Ybar;                // stride 16*T (overwritten)
    int ncols;
    int f;
    // the columns' rating lists cut into parts of at most YBAR_PART ratings:
    // part p = {column, first rating, end}; column c's parts are
    // colpart[c] .. colpart[c+1]-1, in order; partial: one row per part
    const int4 *parts;
    const int *colpart;
    float *partial;             // nparts x 16*T
    int nparts;
};
constexpr int YBAR_PART = 1024;

struct RmseArgs {
    const int *row, *col;
    const float *val;
    const float *U, *V;         // stride ld
    int64_t n;
    int f, ld;
    double *partials;           // one per block
};

// Launchers (als_kernels_*.hip). Each returns the hipError_t of the launch.
int update_lds_bytes(int T, int prec, int solve, int refine);
// variant 0: the default kernel of the solver (two waves per SIMD for the
// refinement solvers); variant 1: one wave per SIMD (warm-started refinement,
// T >= 7; null elsewhere).
hipError_t launch_update(int T, int prec, int solve, int variant, const UpdateArgs &a, int grid, int lds_bytes,
                         hipStream_t s);
const void *update_kernel_ptr(int T, int prec, int solve, int variant = 0);
// Light rows through the dual system, one launch per half-sweep after the
// update kernel (null / 0 where the variant has no dual system: f < 64, CG,
// classic refinement). dual_max_ratings: rows below it take the dual system.
const void *dual_kernel_ptr(int T, int prec, int solve);
int dual_lds_bytes(int T, int prec);
int dual_max_ratings(int T, int prec);
hipError_t launch_dual(int T, int prec, int solve, const UpdateArgs &a, int grid, int lds_bytes, hipStream_t s);
hipError_t launch_adjoint_solve(int T, const AdjointArgs &a, hipStream_t s);
hipError_t launch_cg_batch(int T, const CgBatchArgs &a, hipStream_t s);
hipError_t launch_ybar(int T, const YbarArgs &a, hipStream_t s);
// RMSE partial sums of a.n ratings over `blocks` blocks into a.partials
// (blocks doubles), then out[0] = their sum (with_rmse: out[1] = sqrt(sum / n)).
constexpr int RMSE_BLOCKS = 1024;
hipError_t launch_rmse(const RmseArgs &a, int blocks, double *out, bool with_rmse, hipStream_t s);
hipError_t launch_to16(int prec, const float *src, uint16_t *dst, int64_t count, hipStream_t s);

}  // namespace als
