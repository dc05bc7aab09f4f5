// Instantiates the per-row update kernel for one Gram precision (ALS_PREC is
// set by the build: 0 FP32, 1 BF16, 2 FP16), all tile counts T = 1..9 and
// both solvers. Compiled once per precision so the three build in parallel.
#include "als_device.hpp"

#ifndef ALS_PREC
#error "ALS_PREC must be defined"
#endif

namespace als {

template <int S>
static const void *kernel_w2_for(int T) {
    switch (T) {
    case 1: return (const void *)&update_kernel_w2<1, ALS_PREC, S>;
    case 2: return (const void *)&update_kernel_w2<2, ALS_PREC, S>;
    case 3: return (const void *)&update_kernel_w2<3, ALS_PREC, S>;
    case 4: return (const void *)&update_kernel_w2<4, ALS_PREC, S>;
    case 5: return (const void *)&update_kernel_w2<5, ALS_PREC, S>;
    case 6: return (const void *)&update_kernel_w2<6, ALS_PREC, S>;
    case 7: return (const void *)&update_kernel_w2<7, ALS_PREC, S>;
    case 8: return (const void *)&update_kernel_w2<8, ALS_PREC, S>;
    case 9: return (const void *)&update_kernel_w2<9, ALS_PREC, S>;
    default: return nullptr;
    }
}

// One wave per SIMD (no waves-per-EU bound: the whole register file for one
// wave, spill-free at the widest tile counts); selectable per side.
template <int S>
static const void *kernel_w1_for(int T) {
    switch (T) {
    case 7: return (const void *)&update_kernel<7, ALS_PREC, S>;
    case 8: return (const void *)&update_kernel<8, ALS_PREC, S>;
    case 9: return (const void *)&update_kernel<9, ALS_PREC, S>;
    default: return nullptr;
    }
}

template <int S>
static const void *kernel_for(int T) {
    switch (T) {
    case 1: return (const void *)&update_kernel<1, ALS_PREC, S>;
    case 2: return (const void *)&update_kernel<2, ALS_PREC, S>;
    case 3: return (const void *)&update_kernel<3, ALS_PREC, S>;
    case 4: return (const void *)&update_kernel<4, ALS_PREC, S>;
    case 5: return (const void *)&update_kernel<5, ALS_PREC, S>;
    case 6: return (const void *)&update_kernel<6, ALS_PREC, S>;
    case 7: return (const void *)&update_kernel<7, ALS_PREC, S>;
    case 8: return (const void *)&update_kernel<8, ALS_PREC, S>;
    case 9: return (const void *)&update_kernel<9, ALS_PREC, S>;
    default: return nullptr;
    }
}

template <int T>
static int lds_for(int solve, int refine) {
    (void)refine;
    return solve == SOLVE_CHOL_REFINE ? Geo<T, ALS_PREC>::refine_lds_bytes() : Geo<T, ALS_PREC>::lds_bytes();
}

#define ALS_CAT2(a, b) a##b
#define ALS_CAT(a, b) ALS_CAT2(a, b)

const void *ALS_CAT(update_kernel_ptr_p, ALS_PREC)(int T, int solve, int variant) {
    if (variant == 1) return solve == SOLVE_CHOL_WARM ? kernel_w1_for<SOLVE_CHOL_WARM>(T) : nullptr;
    if (variant != 0) return nullptr;
    return solve == SOLVE_CHOL        ? kernel_for<SOLVE_CHOL>(T)
           : solve == SOLVE_CG        ? kernel_for<SOLVE_CG>(T)
           : solve == SOLVE_CHOL_WARM ? kernel_w2_for<SOLVE_CHOL_WARM>(T)
                                      : kernel_w2_for<SOLVE_CHOL_REFINE>(T);
}

int ALS_CAT(update_lds_bytes_p, ALS_PREC)(int T, int solve, int refine) {
    switch (T) {
    case 1: return lds_for<1>(solve, refine);
    case 2: return lds_for<2>(solve, refine);
    case 3: return lds_for<3>(solve, refine);
    case 4: return lds_for<4>(solve, refine);
    case 5: return lds_for<5>(solve, refine);
    case 6: return lds_for<6>(solve, refine);
    case 7: return lds_for<7>(solve, refine);
    case 8: return lds_for<8>(solve, refine);
    case 9: return lds_for<9>(solve, refine);
    default: return -1;
    }
}

template <int T, int S>
static const void *dual_one() {
    if constexpr (Geo<T, ALS_PREC>::DUAL_TD > 0) return (const void *)&dual_kernel<T, ALS_PREC, S>;
    else return nullptr;
}

template <int S>
static const void *dual_for(int T) {
    switch (T) {
    case 5: return dual_one<5, S>();
    case 6: return dual_one<6, S>();
    case 7: return dual_one<7, S>();
    case 8: return dual_one<8, S>();
    case 9: return dual_one<9, S>();
    default: return nullptr;
    }
}

// The dual system exists for T >= 5 (Geo::DUAL_TD > 0), after the plain
// Cholesky solve or the warm-started one.
const void *ALS_CAT(dual_kernel_ptr_p, ALS_PREC)(int T, int solve) {
    return solve == SOLVE_CHOL        ? dual_for<SOLVE_CHOL>(T)
           : solve == SOLVE_CHOL_WARM ? dual_for<SOLVE_CHOL_WARM>(T)
                                      : nullptr;
}

template <int T>
static int dual_lds_for() {
    using G = Geo<T, ALS_PREC>;
    return G::dual_lds_bytes();
}

int ALS_CAT(dual_lds_bytes_p, ALS_PREC)(int T) {
    switch (T) {
    case 5: return dual_lds_for<5>();
    case 6: return dual_lds_for<6>();
    case 7: return dual_lds_for<7>();
    case 8: return dual_lds_for<8>();
    case 9: return dual_lds_for<9>();
    default: return 0;
    }
}

int ALS_CAT(dual_max_ratings_p, ALS_PREC)(int T) {
    switch (T) {
    case 5: return 16 * Geo<5, ALS_PREC>::DUAL_TD;
    case 6: return 16 * Geo<6, ALS_PREC>::DUAL_TD;
    case 7: return 16 * Geo<7, ALS_PREC>::DUAL_TD;
    case 8: return 16 * Geo<8, ALS_PREC>::DUAL_TD;
    case 9: return 16 * Geo<9, ALS_PREC>::DUAL_TD;
    default: return 0;
    }
}

hipError_t ALS_CAT(launch_update_p, ALS_PREC)(int T, int solve, int variant, const UpdateArgs &a, int grid, int lds,
                                             hipStream_t s) {
    const void *k = ALS_CAT(update_kernel_ptr_p, ALS_PREC)(T, solve, variant);
    if (!k) return hipErrorInvalidValue;
    void *args[] = {const_cast<UpdateArgs *>(&a)};
    return hipLaunchKernel(k, dim3(grid), dim3(WAVE), args, (size_t)lds, s);
}

}  // namespace als
