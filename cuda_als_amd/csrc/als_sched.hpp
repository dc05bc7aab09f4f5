// Work schedule of one half-sweep: rows in decreasing-nnz order (the heavy
// rows start first, so the tail of the launch is made of light rows) cut into
// batches that the persistent waves pop from the device queue.
//
// The reference launches one block per row of a fixed batch in row order
// (als_model.cu:1891-1982) with no load balancing; rows range from 1 to
// about 10^4 ratings at the sizes this engine targets.
#pragma once
#include <cstdint>
#include <vector>

namespace als {

struct Schedule {
    std::vector<int> order;        // row ids, heaviest first (ties by row id)
    std::vector<int> batch_begin;  // batch b = order[batch_begin[b] .. batch_begin[b+1])
};

// Cost model of one row in rating-equivalents: its Gram work plus a fixed
// per-row solve cost that depends on the Gram precision (the solve is FP32
// either way, the Gram is 16x cheaper on BF16/FP16 MFMA).
int64_t row_solve_cost(int prec);

// batch_work <= 0 picks a size that gives each persistent wave about
// 16 batches; a row heavier than the batch size forms a batch on its own.
Schedule build_schedule(int64_t rows, const int *ptr, int batch_work, int grid, int prec);

}  // namespace als
