#include "als_sched.hpp"

#include <algorithm>
#include <numeric>

#include "als_launch.hpp"

namespace als {

int64_t row_solve_cost(int prec) { return prec == PREC_FP32 ? 96 : 1024; }

Schedule build_schedule(int64_t rows, const int *ptr, int batch_work, int grid, int prec) {
    Schedule s;
    s.order.resize(rows);
    std::iota(s.order.begin(), s.order.end(), 0);
    std::stable_sort(s.order.begin(), s.order.end(), [&](int a, int b) {
        return (ptr[a + 1] - ptr[a]) > (ptr[b + 1] - ptr[b]);
    });
    const int64_t solve = row_solve_cost(prec);
    int64_t total = 0;
    for (int64_t r = 0; r < rows; r++) total += (int64_t)(ptr[r + 1] - ptr[r]) + solve;
    int64_t W = batch_work > 0 ? batch_work : std::max<int64_t>(solve, total / std::max<int64_t>(1, (int64_t)grid * 16));
    s.batch_begin.push_back(0);
    int64_t acc = 0;
    for (int64_t u = 0; u < rows; u++) {
        const int r = s.order[u];
        acc += (int64_t)(ptr[r + 1] - ptr[r]) + solve;
        if (acc >= W) {
            s.batch_begin.push_back((int)(u + 1));
            acc = 0;
        }
    }
    if (s.batch_begin.back() != (int)rows) s.batch_begin.push_back((int)rows);
    return s;
}

}  // namespace als
