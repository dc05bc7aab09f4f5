#include "als_sched.hpp"

#include <algorithm>
#include <numeric>

#include "als_launch.hpp"

namespace als {

int64_t row_solve_cost(int prec) { return prec == PREC_FP32 ? 96 : 1024; }

Schedule build_schedule(int64_t rows, const int *ptr, int64_t split_len, int64_t batch_work, int grid, int prec,
                        const std::vector<int> *subset) {
    Schedule s;
    const int64_t solve = row_solve_cost(prec);
    const int64_t nrows = subset ? (int64_t)subset->size() : rows;
    auto row_at = [&](int64_t i) -> int64_t { return subset ? (int64_t)(*subset)[i] : i; };
    int64_t total = 0;
    for (int64_t i = 0; i < nrows; i++) {
        const int64_t r = row_at(i);
        total += (int64_t)(ptr[r + 1] - ptr[r]) + solve;
    }
    const int64_t waves = std::max<int64_t>(1, grid);
    // Parts small enough that the heaviest row spreads over many waves, large
    // enough that the partial-tile traffic stays small next to the gathers.
    s.split_len = split_len > 0 ? split_len : std::max<int64_t>(2048, total / (waves * 8));
    s.batch_work = batch_work > 0 ? batch_work : std::max<int64_t>(solve, total / (waves * 16));

    struct Item { Unit u; int64_t cost; };
    std::vector<Item> items;
    items.reserve(nrows + total / s.split_len + 1);
    for (int64_t i = 0; i < nrows; i++) {
        const int64_t r = row_at(i);
        const int64_t n = ptr[r + 1] - ptr[r];
        if (n <= s.split_len) {
            items.push_back({{(int)r, ptr[r], ptr[r + 1], -1}, n + solve});
            continue;
        }
        const int64_t P = (n + s.split_len - 1) / s.split_len;
        const int64_t len = (n + P - 1) / P;
        Split sp{(int)s.nslots, (int)P, (int)len, 0};
        const int sid = (int)s.splits.size();
        s.splits.push_back(sp);
        s.nslots += P;
        for (int64_t p = 0; p < P; p++) {
            const int64_t kb = ptr[r] + p * len, ke = std::min<int64_t>(ptr[r + 1], kb + len);
            items.push_back({{(int)r, (int)kb, (int)ke, sid}, (ke - kb) + (p == P - 1 ? solve : 0)});
        }
    }
    std::stable_sort(items.begin(), items.end(), [](const Item &a, const Item &b) { return a.cost > b.cost; });
    s.units.resize(items.size());
    s.batch_begin.push_back(0);
    int64_t acc = 0;
    for (size_t i = 0; i < items.size(); i++) {
        s.units[i] = items[i].u;
        acc += items[i].cost;
        if (acc >= s.batch_work) {
            s.batch_begin.push_back((int)(i + 1));
            acc = 0;
        }
    }
    if (s.batch_begin.back() != (int)items.size()) s.batch_begin.push_back((int)items.size());
    return s;
}

}  // namespace als
