// Kernels around the per-row update: test/train RMSE with an FP64 sum,
// FP32 -> 16-bit factor shadows, and the two adjoint kernels; plus the
// dispatch from (T, precision, solver) to the per-precision update objects.
#include "als_device.hpp"

namespace als {

const void *update_kernel_ptr_p0(int T, int solve, int variant);
const void *update_kernel_ptr_p1(int T, int solve, int variant);
const void *update_kernel_ptr_p2(int T, int solve, int variant);
int update_lds_bytes_p0(int T, int solve, int refine);
int update_lds_bytes_p1(int T, int solve, int refine);
int update_lds_bytes_p2(int T, int solve, int refine);
hipError_t launch_update_p0(int T, int solve, int variant, const UpdateArgs &a, int grid, int lds, hipStream_t s);
hipError_t launch_update_p1(int T, int solve, int variant, const UpdateArgs &a, int grid, int lds, hipStream_t s);
hipError_t launch_update_p2(int T, int solve, int variant, const UpdateArgs &a, int grid, int lds, hipStream_t s);

const void *update_kernel_ptr(int T, int prec, int solve, int variant) {
    switch (prec) {
    case PREC_FP32: return update_kernel_ptr_p0(T, solve, variant);
    case PREC_BF16: return update_kernel_ptr_p1(T, solve, variant);
    case PREC_FP16: return update_kernel_ptr_p2(T, solve, variant);
    default: return nullptr;
    }
}

int update_lds_bytes(int T, int prec, int solve, int refine) {
    switch (prec) {
    case PREC_FP32: return update_lds_bytes_p0(T, solve, refine);
    case PREC_BF16: return update_lds_bytes_p1(T, solve, refine);
    case PREC_FP16: return update_lds_bytes_p2(T, solve, refine);
    default: return -1;
    }
}

hipError_t launch_update(int T, int prec, int solve, int variant, const UpdateArgs &a, int grid, int lds,
                         hipStream_t s) {
    switch (prec) {
    case PREC_FP32: return launch_update_p0(T, solve, variant, a, grid, lds, s);
    case PREC_BF16: return launch_update_p1(T, solve, variant, a, grid, lds, s);
    case PREC_FP16: return launch_update_p2(T, solve, variant, a, grid, lds, s);
    default: return hipErrorInvalidValue;
    }
}

const void *dual_kernel_ptr_p0(int T, int solve);
const void *dual_kernel_ptr_p1(int T, int solve);
const void *dual_kernel_ptr_p2(int T, int solve);
int dual_lds_bytes_p0(int T);
int dual_lds_bytes_p1(int T);
int dual_lds_bytes_p2(int T);
int dual_max_ratings_p0(int T);
int dual_max_ratings_p1(int T);
int dual_max_ratings_p2(int T);

const void *dual_kernel_ptr(int T, int prec, int solve) {
    switch (prec) {
    case PREC_FP32: return dual_kernel_ptr_p0(T, solve);
    case PREC_BF16: return dual_kernel_ptr_p1(T, solve);
    case PREC_FP16: return dual_kernel_ptr_p2(T, solve);
    default: return nullptr;
    }
}

int dual_lds_bytes(int T, int prec) {
    switch (prec) {
    case PREC_FP32: return dual_lds_bytes_p0(T);
    case PREC_BF16: return dual_lds_bytes_p1(T);
    case PREC_FP16: return dual_lds_bytes_p2(T);
    default: return 0;
    }
}

int dual_max_ratings(int T, int prec) {
    switch (prec) {
    case PREC_FP32: return dual_max_ratings_p0(T);
    case PREC_BF16: return dual_max_ratings_p1(T);
    case PREC_FP16: return dual_max_ratings_p2(T);
    default: return 0;
    }
}

hipError_t launch_dual(int T, int prec, int solve, const UpdateArgs &a, int grid, int lds, hipStream_t s) {
    const void *k = dual_kernel_ptr(T, prec, solve);
    if (!k) return hipErrorInvalidValue;
    void *args[] = {const_cast<UpdateArgs *>(&a)};
    return hipLaunchKernel(k, dim3(grid), dim3(WAVE), args, (size_t)lds, s);
}

// ------------------------------------------------------------------ RMSE
// Reference: calculate_square_error (als_model.cu:62-89) + cublasSasum +
// host sqrt (:1744-1764). Sixteen lanes share one rating: each loads float4
// slices of u and v, the FP32 dot is reduced across the 16 lanes, and the
// squared error is accumulated in FP64. A lane group takes RMSE_ILP of its
// ratings at once (every index and row load issued before the first dot),
// so that their memory latencies overlap, and adds their errors in rating
// order. Every block writes one partial; a second kernel sums the partials
// in a fixed order, so the result is deterministic (no float atomics).
constexpr int RMSE_THREADS = 256;
constexpr int RMSE_ILP = 4;

__global__ __launch_bounds__(RMSE_THREADS) void rmse_kernel(RmseArgs a) {
    __shared__ double red[RMSE_THREADS / WAVE];
    const int sub = threadIdx.x & 15;
    const int64_t group = ((int64_t)blockIdx.x * RMSE_THREADS + threadIdx.x) >> 4;
    const int64_t ngroups = (int64_t)gridDim.x * (RMSE_THREADS / 16);
    double acc = 0.0;
    for (int64_t t0 = group; t0 < a.n; t0 += RMSE_ILP * ngroups) {
        const float *u[RMSE_ILP], *v[RMSE_ILP];
        float r[RMSE_ILP], d[RMSE_ILP];
        bool in[RMSE_ILP];
#pragma unroll
        for (int i = 0; i < RMSE_ILP; ++i) {
            const int64_t t = t0 + i * ngroups;
            in[i] = t < a.n;
            const int64_t tc = in[i] ? t : t0;  // a valid rating, not added
            u[i] = a.U + (int64_t)a.row[tc] * a.ld;
            v[i] = a.V + (int64_t)a.col[tc] * a.ld;
            r[i] = a.val[tc];
            d[i] = 0.f;
        }
        for (int j = 4 * sub; j < a.f; j += 64) {
#pragma unroll
            for (int i = 0; i < RMSE_ILP; ++i) {
                const f32x4 x = *reinterpret_cast<const f32x4 *>(u[i] + j);
                const f32x4 y = *reinterpret_cast<const f32x4 *>(v[i] + j);
                // columns f.. of the last slice are masked (padding of a caller's rows need not be zero)
                const float p3 = j + 3 < a.f ? x[3] * y[3] : 0.f, p2 = j + 2 < a.f ? x[2] * y[2] : 0.f;
                const float p1 = j + 1 < a.f ? x[1] * y[1] : 0.f;
                d[i] += x[0] * y[0] + p1 + p2 + p3;
            }
        }
#pragma unroll
        for (int i = 0; i < RMSE_ILP; ++i) {
#pragma unroll
            for (int m = 8; m >= 1; m >>= 1) d[i] += __shfl_xor(d[i], m, 16);
        }
        if (sub == 0) {
#pragma unroll
            for (int i = 0; i < RMSE_ILP; ++i) {
                const double e = (double)r[i] - (double)d[i];
                if (in[i]) acc += e * e;
            }
        }
    }
    acc = wave_sum_d(acc);
    if ((threadIdx.x & (WAVE - 1)) == 0) red[threadIdx.x / WAVE] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < RMSE_THREADS / WAVE; ++w) s += red[w];
        a.partials[blockIdx.x] = s;
    }
}

// out[0] = the sum of the partials; with with_rmse, out[1] = sqrt(sum / n).
__global__ __launch_bounds__(WAVE) void sum_partials_kernel(const double *p, int nb, double *out, int64_t n,
                                                            int with_rmse) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += WAVE) s += p[i];
    s = wave_sum_d(s);
    if (threadIdx.x == 0) {
        out[0] = s;
        if (with_rmse) out[1] = sqrt(s / (double)n);
    }
}

hipError_t launch_rmse(const RmseArgs &a, int blocks, double *out, bool with_rmse, hipStream_t s) {
    rmse_kernel<<<blocks, RMSE_THREADS, 0, s>>>(a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    sum_partials_kernel<<<1, WAVE, 0, s>>>(a.partials, blocks, out, a.n, with_rmse ? 1 : 0);
    return hipGetLastError();
}

// ---------------------------------------------------------- 16-bit shadow
// The reference converts the fixed factors to FP16 before every tensor-core
// Gram (float2half_array, als_model.cu:54-60, 1883-1889). Here the update
// kernel refreshes the shadow as it writes each row; this kernel builds it
// after factors are set from the host.
template <int E>
__global__ void to16_kernel(const float *src, uint16_t *dst, int64_t count) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = to16<E>(src[i]);
}

hipError_t launch_to16(int prec, const float *src, uint16_t *dst, int64_t count, hipStream_t s) {
    const int blocks = (int)std::min<int64_t>((count + 255) / 256, 65536);
    if (count == 0) return hipSuccess;
    if (prec == PREC_BF16) to16_kernel<PREC_BF16><<<blocks, 256, 0, s>>>(src, dst, count);
    else if (prec == PREC_FP16) to16_kernel<PREC_FP16><<<blocks, 256, 0, s>>>(src, dst, count);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

// --------------------------------------------------------------- adjoint
// No reference counterpart. For x_u = A_u^-1 b_u and an upstream gradient
// g_u = dL/dx_u, with the Cholesky factor saved by the forward pass:
//   xbar_u = A_u^-1 g_u  (forward + back substitution on the saved factor)
//   dL/dlambda_u = -n_u xbar_u . x_u
template <int T>
__global__ __launch_bounds__(WAVE) void adjoint_solve_kernel(AdjointArgs a) {
    constexpr int FP = 16 * T, NT = T * (T + 1) / 2;
    const int l = lane(), f = a.f;
    for (int row = blockIdx.x; row < a.rows; row += gridDim.x) {
        const int n = a.ptr[row + 1] - a.ptr[row];
        const int64_t o = (int64_t)row * FP;
        float g[T];
#pragma unroll
        for (int J = 0; J < T; ++J) g[J] = 0.f;
        float lb = 0.f;
        if (n > 0) {
            f32x4 L[NT];
            const float *src = a.Lsave + (int64_t)row * NT * 256;
#pragma unroll
            for (int t = 0; t < NT; ++t) L[t] = *reinterpret_cast<const f32x4 *>(src + t * 256 + 4 * l);
            load_colform<T>(a.Gbar + o, f, g);
            f32x4 wr[T];
            fwd_solve<T>(L, f, g, wr);
            back_solve<T>(L, f, wr, g);
            float x[T];
            load_colform<T>(a.X + o, f, x);
            lb = -(float)n * colform_dot(g, x, T);
        }
        store_colform<T>(a.Xbar + o, g);
        if (l == 0) a.lambda_bar[row] = lb;
    }
}

// dL/dy_i = sum_{u in Omega(i)} r_ui xbar_u - xbar_u (x_u . y_i) - x_u (xbar_u . y_i)
// Two passes, no atomics, a fixed summation order. ybar_part_kernel: one
// wave per part of a column's rating list (at most YBAR_PART ratings, so a
// popular column is spread over many waves); its four 16-lane groups take
// every fourth rating, lane c of a group holding factors c, c + 16, ...,
// c + 16(T-1); the two dots are 16-lane sums; the groups' sums are added at
// the end and written as the part's row. ybar_sum_kernel: one wave per
// column adds its parts' rows in part order.
template <int T>
__global__ __launch_bounds__(WAVE) void ybar_part_kernel(YbarArgs a) {
    constexpr int FP = 16 * T;
    const int l = lane(), c = l & 15, grp = l >> 4, f = a.f;
    for (int p = blockIdx.x; p < a.nparts; p += gridDim.x) {
        const int4 pt = a.parts[p];
        const int64_t oy = (int64_t)pt.x * FP;
        float y[T], ac[T];
#pragma unroll
        for (int J = 0; J < T; ++J) {
            const int j = 16 * J + c;
            y[J] = j < f ? a.Y[oy + j] : 0.f;
            ac[J] = 0.f;
        }
        for (int k = pt.y + grp; k < pt.z; k += 4) {
            const int64_t ox = (int64_t)a.cidx[k] * FP;
            const float r = a.cval[k];
            float x[T], b[T], pxy = 0.f, pby = 0.f;
#pragma unroll
            for (int J = 0; J < T; ++J) {
                const int j = 16 * J + c;
                x[J] = j < f ? a.X[ox + j] : 0.f;
                b[J] = j < f ? a.Xbar[ox + j] : 0.f;
                pxy += x[J] * y[J];
                pby += b[J] * y[J];
            }
#pragma unroll
            for (int m = 8; m >= 1; m >>= 1) {
                pxy += __shfl_xor(pxy, m, 16);
                pby += __shfl_xor(pby, m, 16);
            }
#pragma unroll
            for (int J = 0; J < T; ++J) ac[J] += r * b[J] - b[J] * pxy - x[J] * pby;
        }
#pragma unroll
        for (int J = 0; J < T; ++J) {
            ac[J] += __shfl_xor(ac[J], 16, WAVE);
            ac[J] += __shfl_xor(ac[J], 32, WAVE);
        }
        if (grp == 0) {
#pragma unroll
            for (int J = 0; J < T; ++J) a.partial[(int64_t)p * FP + 16 * J + c] = ac[J];
        }
    }
}

template <int T>
__global__ __launch_bounds__(WAVE) void ybar_sum_kernel(YbarArgs a) {
    constexpr int FP = 16 * T, NJ = (FP + WAVE - 1) / WAVE;
    const int l = lane(), f = a.f;
    for (int col = blockIdx.x; col < a.ncols; col += gridDim.x) {
        float s[NJ];
#pragma unroll
        for (int m = 0; m < NJ; ++m) s[m] = 0.f;
        for (int p = a.colpart[col]; p < a.colpart[col + 1]; ++p) {
#pragma unroll
            for (int m = 0; m < NJ; ++m) {
                const int j = WAVE * m + l;
                if (j < FP) s[m] += a.partial[(int64_t)p * FP + j];
            }
        }
#pragma unroll
        for (int m = 0; m < NJ; ++m) {
            const int j = WAVE * m + l;
            if (j < FP) a.Ybar[(int64_t)col * FP + j] = j < f ? s[m] : 0.f;
        }
    }
}

// One wave per system: the lower 16x16 tiles of A in the register layout
// of the update kernel (diagonal tiles A[16I+4q+r][16I+c], off-diagonal
// tiles row by lane A[16I+c][16J+4q+r]), zero outside the f x f block,
// then the same register CG as the fused update (cumf recurrence: warm
// start, cg_iters cap, exit once r.r < tol after an update, pAp == 0 guard).
// Up to 7 tiles two waves per SIMD (256 VGPRs; at 7 tiles a few loop-
// invariant values spill, none inside the iterations), so that one
// system's loads and reduction chains run beside another's; 8 and 9 tiles
// keep the whole register file.
template <int T>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(T <= 7 ? 2 : 1))) void cg_batch_kernel(
    CgBatchArgs a) {
    constexpr int NT = T * (T + 1) / 2;
    const int l = lane(), q = l >> 4, c = l & 15;
    const int f = a.f;
    for (int sys = blockIdx.x; sys < a.batch; sys += gridDim.x) {
        const int64_t base = (int64_t)sys * f * f;
        auto at = [&](int i, int j) -> float {
            if (i >= f || j >= f) return 0.f;
            const int64_t k = base + (int64_t)i * f + j;
            if (a.A16) return (float)__builtin_bit_cast(_Float16, a.A16[k]);
            if (a.round_a16) return (float)(_Float16)a.A[k];
            return a.A[k];
        };
        f32x4 acc[NT];
        if (a.vec) {
            // lane (q, c) of tile (I, J): row 16I + c, columns 16J + 4q .. +3,
            // one load (a diagonal tile's entries A[16I+4q+r][16I+c] read as
            // their mirror A[16I+c][16I+4q+r], A being symmetric); with
            // f % 4 == 0 the four columns are all inside f or all past it
            // Every tile's load is issued before any value is converted:
            // unconditional buffer loads through a descriptor over this
            // system's A, a tile past f reading zeros by the range check (a
            // bounds test around each load would hold its conversion and
            // wait for the load there, one memory round trip per tile). An
            // FP16 tile lands in the low half of its accumulator's bits and
            // is widened in place afterwards.
            typedef _Float16 f16x4v __attribute__((ext_vector_type(4)));
            const int esz = a.A16 ? 2 : 4;
            const void *sysA = a.A16 ? (const void *)(a.A16 + base) : (const void *)(a.A + base);
            const __amdgpu_buffer_rsrc_t rsA =
                __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(sysA), (short)0, f * f * esz, 0x00020000);
            auto off = [&](int I, int J) {
                const int i = 16 * I + c, j = 16 * J + 4 * q;
                return (i < f && j < f) ? (i * f + j) * esz : 0x40000000;  // f * f * esz < 2^17
            };
            if (a.A16) {
#pragma unroll
                for (int I = 0; I < T; ++I)
#pragma unroll
                    for (int J = 0; J <= I; ++J) {
                        const u32x2 w = __builtin_bit_cast(u32x2, __builtin_amdgcn_raw_buffer_load_b64(rsA, off(I, J), 0, 0));
                        acc[tile_id(I, J)] = __builtin_bit_cast(f32x4, u32x4{w[0], w[1], 0u, 0u});
                    }
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const u32x4 u = __builtin_bit_cast(u32x4, acc[t]);
                    acc[t] = __builtin_convertvector(__builtin_bit_cast(f16x4v, u32x2{u[0], u[1]}), f32x4);
                }
            } else {
#pragma unroll
                for (int I = 0; I < T; ++I)
#pragma unroll
                    for (int J = 0; J <= I; ++J)
                        acc[tile_id(I, J)] =
                            __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(rsA, off(I, J), 0, 0));
                if (a.round_a16) {
#pragma unroll
                    for (int t = 0; t < NT; ++t)
                        acc[t] = __builtin_convertvector(__builtin_convertvector(acc[t], f16x4v), f32x4);
                }
            }
        } else {
#pragma unroll
            for (int I = 0; I < T; ++I)
#pragma unroll
                for (int J = 0; J <= I; ++J)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        acc[tile_id(I, J)][r] = (I == J) ? at(16 * I + 4 * q + r, 16 * I + c)
                                                         : at(16 * I + c, 16 * J + 4 * q + r);
        }
        float bc[T], xc[T];
#pragma unroll
        for (int J = 0; J < T; ++J) {
            const int i = 16 * J + c;
            bc[J] = i < f ? a.b[(int64_t)sys * f + i] : 0.f;
            xc[J] = i < f ? a.x[(int64_t)sys * f + i] : 0.f;
        }
        cg_regs<T>(acc, f, bc, xc, a.iters, a.tol);
        if (q == 0) {
#pragma unroll
            for (int J = 0; J < T; ++J)
                if (16 * J + c < f) a.x[(int64_t)sys * f + 16 * J + c] = xc[J];
        }
    }
}

hipError_t launch_cg_batch(int T, const CgBatchArgs &a, hipStream_t s) {
    const int grid = std::max(1, std::min(a.batch, 65536));
    switch (T) {
#define ALS_CGB(TT) case TT: cg_batch_kernel<TT><<<grid, WAVE, 0, s>>>(a); break;
    ALS_CGB(1) ALS_CGB(2) ALS_CGB(3) ALS_CGB(4) ALS_CGB(5) ALS_CGB(6) ALS_CGB(7) ALS_CGB(8) ALS_CGB(9)
#undef ALS_CGB
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_adjoint_solve(int T, const AdjointArgs &a, hipStream_t s) {
    const int grid = std::max(1, std::min(a.rows, 65536));
    const size_t lds = 0;
    switch (T) {
#define ALS_ADJ(TT) case TT: adjoint_solve_kernel<TT><<<grid, WAVE, lds, s>>>(a); break;
    ALS_ADJ(1) ALS_ADJ(2) ALS_ADJ(3) ALS_ADJ(4) ALS_ADJ(5) ALS_ADJ(6) ALS_ADJ(7) ALS_ADJ(8) ALS_ADJ(9)
#undef ALS_ADJ
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_ybar(int T, const YbarArgs &a, hipStream_t s) {
    const int gp = std::max(1, std::min(a.nparts, 65536)), gc = std::max(1, std::min(a.ncols, 65536));
    switch (T) {
#define ALS_YB(TT)                                                   \
    case TT:                                                         \
        ybar_part_kernel<TT><<<gp, WAVE, 0, s>>>(a);                 \
        ybar_sum_kernel<TT><<<gc, WAVE, 0, s>>>(a);                  \
        break;
    ALS_YB(1) ALS_YB(2) ALS_YB(3) ALS_YB(4) ALS_YB(5) ALS_YB(6) ALS_YB(7) ALS_YB(8) ALS_YB(9)
#undef ALS_YB
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace als
