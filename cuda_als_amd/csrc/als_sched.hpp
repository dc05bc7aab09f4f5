// Work schedule of one half-sweep.
//
// The reference launches one block per row of a fixed batch in row order
// (als_model.cu:1891-1982) with no load balancing; at the sizes this engine
// targets a row has from 1 to ~5*10^5 ratings. Here:
//   * a row heavier than `split_len` ratings is cut into parts of near-equal
//     length; every part is its own work unit, each writes its partial Gram
//     tiles to scratch and the last part to finish sums the partials in part
//     order (deterministic) and solves;
//   * units are ordered by decreasing cost (longest-processing-time first),
//     so the tail of the launch is made of the lightest rows;
//   * consecutive units are grouped into batches of about `batch_work`
//     rating-equivalents; persistent waves pop batches from a device queue.
#pragma once
#include <cstdint>
#include <vector>

namespace als {

struct Unit {
    int row;      // row of the side being updated
    int kb, ke;   // rating range [kb, ke) inside the row's CSR segment (absolute)
    int split;    // index into Schedule::splits, or -1 for a whole row
};

struct Split {
    int slot_base;  // first partial-Gram slot of this row
    int nparts;
    int part_len;   // every part but the last has exactly this many ratings
    int pad;
};

struct Schedule {
    std::vector<Unit> units;
    std::vector<Split> splits;
    std::vector<int> batch_begin;  // batch b = units[batch_begin[b] .. batch_begin[b+1])
    int64_t nslots = 0;            // partial-Gram slots needed
    int64_t split_len = 0;
    int64_t batch_work = 0;
};

// Cost of a solve in rating-equivalents of Gram work for the given Gram
// precision (the solve is FP32 either way; the 16-bit Gram is ~16x cheaper
// per rating than the FP32 one).
int64_t row_solve_cost(int prec);

// split_len <= 0 and batch_work <= 0 select sizes from the total work and
// the number of persistent waves (`grid`). subset (nullable): schedule only
// these rows (ascending), e.g. the rows left after the dual system's.
Schedule build_schedule(int64_t rows, const int *ptr, int64_t split_len, int64_t batch_work, int grid, int prec,
                        const std::vector<int> *subset = nullptr);

}  // namespace als
