// One persistent launch for several models of one rating matrix (different
// factor ranks f, so different tile counts T): the waves pop batches from a
// combined queue and run each with the code of that model's T. Compiled once
// per Gram precision (ALS_PREC set by the build), for the Cholesky solve
// without refinement and T = 2..9 (f >= 16): with the single-block body, or
// the refinement code, in the same function the compiler takes many minutes.
#include "als_device.hpp"

#ifndef ALS_PREC
#error "ALS_PREC must be defined"
#endif

namespace als {

// One function for every T: its register allocation is set by the largest
// path (T = 9 needs more than half the VGPR file), so the kernel is held to
// two waves per SIMD; the large-T paths then spill, but every model runs at
// two waves instead of one (f = 32/64/128 at the netflix shape: 29.8 ms per
// sweep against 43.8 ms unhinted; three separate engines take 26.4 ms).
template <int E, int S>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(2, 2))) void update_group_kernel(GroupArgs g) {
    extern __shared__ __attribute__((aligned(16))) char lds[];
    const int total = g.first[g.nmodels];
    for (;;) {
        const int b = pop_batch(g.queue);
        if (b >= total) break;
        int k = 0;
        while (k + 1 < g.nmodels && b >= g.first[k + 1]) ++k;
        const int lb = b - g.first[k];
        switch (g.T[k]) {
        case 2: run_batch<2, E, S, true>(g.m[k], lb, lds); break;
        case 3: run_batch<3, E, S, true>(g.m[k], lb, lds); break;
        case 4: run_batch<4, E, S, true>(g.m[k], lb, lds); break;
        case 5: run_batch<5, E, S, true>(g.m[k], lb, lds); break;
        case 6: run_batch<6, E, S, true>(g.m[k], lb, lds); break;
        case 7: run_batch<7, E, S, true>(g.m[k], lb, lds); break;
        case 8: run_batch<8, E, S, true>(g.m[k], lb, lds); break;
        case 9: run_batch<9, E, S, true>(g.m[k], lb, lds); break;
        default: break;
        }
    }
    leave_queue(g.queue);
}

#define ALS_CAT2(a, b) a##b
#define ALS_CAT(a, b) ALS_CAT2(a, b)

const void *ALS_CAT(update_group_kernel_ptr_p, ALS_PREC)(int solve) {
    return solve == SOLVE_CHOL ? (const void *)&update_group_kernel<ALS_PREC, SOLVE_CHOL> : nullptr;
}

hipError_t ALS_CAT(launch_update_group_p, ALS_PREC)(int solve, const GroupArgs &g, int grid, int lds,
                                                   hipStream_t s) {
    const void *k = ALS_CAT(update_group_kernel_ptr_p, ALS_PREC)(solve);
    if (!k) return hipErrorInvalidValue;
    void *args[] = {const_cast<GroupArgs *>(&g)};
    return hipLaunchKernel(k, dim3(grid), dim3(WAVE), args, (size_t)lds, s);
}

}  // namespace als
