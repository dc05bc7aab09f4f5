// Device code of the MI355X ALS engine (gfx950, wave64, MFMA).
//
// One wavefront owns one rating row at a time (a user in the user
// half-sweep, an item in the item half-sweep) and runs the whole per-row
// update of the reference's update-U/update-V blocks (als_model.cu:1797-2029,
// 2030-2267) without leaving the chip:
//
//   1. Gather the rated rows y_i of the fixed factor matrix, 32 ratings at a
//      time, into LDS, with the rating r_i written into the spare column f
//      ("augmented" row [y_i | r_i]).
//   2. Accumulate the lower-triangle 16x16 tiles of G = sum [y r]^T [y r] on
//      MFMA (FP32 16x16x4, or BF16/FP16 16x16x32 with FP32 accumulators).
//      G[0:f,0:f] is the Gram sum y y^T (reference als_model.cu:91-167),
//      G[f,0:f] is b = sum r y (reference RHS, als_model.cu:1816-1852).
//   3. Add lambda * n_u on the diagonal (als_model.cu:156-159).
//   4. Solve: Cholesky of the augmented matrix in the MFMA accumulator
//      registers (row f of the factor is z = L^-1 b, so the forward
//      substitution is part of the factorisation), then back substitution,
//      optional FP64-residual iterative refinement; or the cuMF CG
//      recurrence (cumf_als_cg/cg.cu:37-232).
//   5. Write x_u in FP32 and, for the mixed-precision path, its 16-bit
//      shadow that the next half-sweep gathers.
//
// Waves pull batches of rows from a device work queue (rows pre-sorted by
// nnz, heavy first); the last wave to leave resets the queue, so the launch
// can be replayed from a graph without any memset.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "als_launch.hpp"

namespace als {

constexpr int WAVE = 64;
constexpr int CHUNK = 32;

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ constexpr int tile_id(int I, int J) { return I * (I + 1) / 2 + J; }

__device__ __forceinline__ int lane() { return (int)threadIdx.x; }

__device__ __forceinline__ float bcast_lane(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

__device__ __forceinline__ float pick4(const f32x4 &v, int i) {
    return i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : v[3];
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, WAVE);
    return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, WAVE);
    return v;
}

template <int E> __device__ __forceinline__ uint16_t to16(float x);
template <> __device__ __forceinline__ uint16_t to16<PREC_BF16>(float x) {
    return __builtin_bit_cast(uint16_t, (__bf16)x);
}
template <> __device__ __forceinline__ uint16_t to16<PREC_FP16>(float x) {
    return __builtin_bit_cast(uint16_t, (_Float16)x);
}
template <int E> __device__ __forceinline__ float from16(uint16_t h);
template <> __device__ __forceinline__ float from16<PREC_BF16>(uint16_t h) {
    return (float)__builtin_bit_cast(__bf16, h);
}
template <> __device__ __forceinline__ float from16<PREC_FP16>(uint16_t h) {
    return (float)__builtin_bit_cast(_Float16, h);
}

// LDS carve-up of one wave (the workgroup is exactly one wave).
template <int T, int E> struct Geo {
    static constexpr int NT = T * (T + 1) / 2;  // lower-triangle 16x16 tiles
    static constexpr int FP = 16 * T;           // padded augmented dimension
    static constexpr int YT_H = 40;             // 16-bit stage row: 32 ratings + 8 pad (80 B)
    static constexpr int YR = (T & 1) ? FP : FP + 16;  // FP32 stage row stride (floats)
    static constexpr int STAGE_BYTES = E == PREC_FP32 ? CHUNK * YR * 4 : FP * YT_H * 2;
    static constexpr int PANEL_BYTES = (T - 1) * 16 * 17 * 4;  // Cholesky panel tiles (after the Gram)
    static constexpr int MISC_FLOATS = 32 + 32 + 128 + 128;  // ibuf, rbuf, colbuf, xbuf
    static constexpr int R1_BYTES = STAGE_BYTES > PANEL_BYTES ? STAGE_BYTES : PANEL_BYTES;
    __host__ __device__ static constexpr int lds_bytes() { return MISC_FLOATS * 4 + R1_BYTES; }
};

// ------------------------------------------------------------------ Gram
// Accumulates the lower tiles of the augmented Gram of one row into acc.
// acc[tile_id(I,J)][r] holds G[16I + 4q + r][16J + c] (q = lane/16,
// c = lane%16), the MFMA 16x16 accumulator layout.
template <int T, int E>
__device__ __forceinline__ void gram_row(const UpdateArgs &a, int64_t rs, int n,
                                         f32x4 (&acc)[Geo<T, E>::NT], char *lds) {
    using G = Geo<T, E>;
    const int l = lane();
    const int f = a.f;
    int *ibuf = reinterpret_cast<int *>(lds);
    float *rbuf = reinterpret_cast<float *>(lds) + 32;
    char *stage = lds + G::MISC_FLOATS * 4;

#pragma unroll
    for (int t = 0; t < G::NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};

    if constexpr (E == PREC_FP32) {
        float *Ys = reinterpret_cast<float *>(stage);
        const int nseg = (f + 3) >> 2;          // float4 segments carrying data
        // Columns past the loaded segments are zero (the solve reuses this
        // region, so clear them for every row).
        for (int e = l; e < CHUNK * (G::FP - 4 * nseg); e += WAVE) {
            int k = e / (G::FP - 4 * nseg), col = 4 * nseg + e % (G::FP - 4 * nseg);
            Ys[k * G::YR + col] = 0.f;
        }
        const int dk = WAVE / nseg, ds = WAVE % nseg;
        for (int k0 = 0; k0 < n; k0 += CHUNK) {
            const int nv = min(CHUNK, n - k0);
            if (l < CHUNK) {
                ibuf[l] = l < nv ? a.idx[rs + k0 + l] : -1;
                rbuf[l] = l < nv ? a.val[rs + k0 + l] : 0.f;
            }
            __syncthreads();
            int k = l / nseg, s = l % nseg;
            for (int task = l; task < CHUNK * nseg; task += WAVE) {
                const int id = ibuf[k];
                const float *src = id >= 0 ? a.Y + (int64_t)id * G::FP + 4 * s : a.zero32 + 4 * s;
                const f32x4 v = *reinterpret_cast<const f32x4 *>(src);
                *reinterpret_cast<f32x4 *>(Ys + k * G::YR + 4 * s) = v;
                k += dk;
                s += ds;
                if (s >= nseg) { s -= nseg; ++k; }
            }
            __syncthreads();
            if (l < CHUNK) Ys[l * G::YR + f] = rbuf[l];
            __syncthreads();
#pragma unroll
            for (int ks = 0; ks < CHUNK / 4; ++ks) {
                float F[T];
#pragma unroll
                for (int I = 0; I < T; ++I) F[I] = Ys[(4 * ks + (l >> 4)) * G::YR + 16 * I + (l & 15)];
#pragma unroll
                for (int I = 0; I < T; ++I)
#pragma unroll
                    for (int J = 0; J <= I; ++J)
                        acc[tile_id(I, J)] =
                            __builtin_amdgcn_mfma_f32_16x16x4f32(F[I], F[J], acc[tile_id(I, J)], 0, 0, 0);
            }
            __syncthreads();
        }
    } else {
        uint16_t *Yt = reinterpret_cast<uint16_t *>(stage);  // Yt[factor][rating]
        const uint16_t *Ysh = a.Ys;
        const int nseg = (f + 7) >> 3;          // 16-byte segments carrying data
        for (int e = l; e < (G::FP - 8 * nseg) * CHUNK; e += WAVE) {
            int j = 8 * nseg + e / CHUNK, k = e % CHUNK;
            Yt[j * G::YT_H + k] = 0;
        }
        for (int k0 = 0; k0 < n; k0 += CHUNK) {
            const int nv = min(CHUNK, n - k0);
            if (l < CHUNK) {
                ibuf[l] = l < nv ? a.idx[rs + k0 + l] : -1;
                rbuf[l] = l < nv ? a.val[rs + k0 + l] : 0.f;
            }
            __syncthreads();
            // Task = (group of 8 ratings g, segment s of 8 factors): eight
            // 16-byte loads, an 8x8 transpose of 16-bit values in registers,
            // eight 16-byte LDS writes of Yt[8s+e][8g..8g+7].
            for (int task = l; task < 4 * nseg; task += WAVE) {
                const int g = task / nseg, s = task - g * nseg;
                u32x4 d[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    // Unused slots of the last chunk read a row of zeros:
                    // the load itself is unconditional.
                    const int id = ibuf[8 * g + e];
                    const uint16_t *src = id >= 0 ? Ysh + (int64_t)id * G::FP + 8 * s : a.zero16 + 8 * s;
                    d[e] = *reinterpret_cast<const u32x4 *>(src);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    u32x4 lo, hi;
#pragma unroll
                    for (int w = 0; w < 4; ++w) {
                        const uint32_t x0 = d[2 * w][v], x1 = d[2 * w + 1][v];
                        lo[w] = __builtin_amdgcn_perm(x1, x0, 0x05040100u);  // low halves of x0, x1
                        hi[w] = __builtin_amdgcn_perm(x1, x0, 0x07060302u);  // high halves
                    }
                    *reinterpret_cast<u32x4 *>(Yt + (8 * s + 2 * v) * G::YT_H + 8 * g) = lo;
                    *reinterpret_cast<u32x4 *>(Yt + (8 * s + 2 * v + 1) * G::YT_H + 8 * g) = hi;
                }
            }
            __syncthreads();
            if (l < CHUNK) Yt[f * G::YT_H + l] = to16<E>(rbuf[l]);
            __syncthreads();
            u32x4 F[T];
#pragma unroll
            for (int I = 0; I < T; ++I)
                F[I] = *reinterpret_cast<const u32x4 *>(Yt + (16 * I + (l & 15)) * G::YT_H + 8 * (l >> 4));
#pragma unroll
            for (int I = 0; I < T; ++I)
#pragma unroll
                for (int J = 0; J <= I; ++J) {
                    if constexpr (E == PREC_BF16)
                        acc[tile_id(I, J)] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            __builtin_bit_cast(bf16x8, F[I]), __builtin_bit_cast(bf16x8, F[J]), acc[tile_id(I, J)],
                            0, 0, 0);
                    else
                        acc[tile_id(I, J)] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                            __builtin_bit_cast(f16x8, F[I]), __builtin_bit_cast(f16x8, F[J]), acc[tile_id(I, J)],
                            0, 0, 0);
                }
            __syncthreads();
        }
    }
}

// Diagonal: + lambda * n on the f real pivots; padding pivots (index > f)
// become 1 so the padded system is the identity there. Index f (the rating
// column's own square sum) is left alone: it is never a pivot.
template <int T, int NT>
__device__ __forceinline__ void add_diag(f32x4 (&acc)[NT], int f, float reg) {
    const int l = lane(), q = l >> 4, c = l & 15;
#pragma unroll
    for (int I = 0; I < T; ++I)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const bool on = 4 * q + r == c;
            const int d = 16 * I + c;
            float v = acc[tile_id(I, I)][r];
            v = (on && d < f) ? v + reg : v;
            v = (on && d > f) ? 1.f : v;
            acc[tile_id(I, I)][r] = v;
        }
}

// ------------------------------------------------------- vector forms
// A vector v of the padded dimension 16T lives in one of two register forms:
//   column form  vc[J]      = v[16J + c]          (replicated over q)
//   row form     vr[I][r]   = v[16I + 4q + r]     (replicated over c)
// matching the columns and the rows a lane owns in the accumulator tiles.

__device__ __forceinline__ f32x4 col_to_row(float vcJ) {
    const int l = lane(), q = l >> 4;
    f32x4 o;
#pragma unroll
    for (int r = 0; r < 4; ++r) o[r] = __shfl(vcJ, (l & 48) | (4 * q + r), WAVE);
    return o;
}

__device__ __forceinline__ float row_to_col(const f32x4 &vrJ) {
    const int c = lane() & 15, src = (c >> 2) << 4;
    const float t0 = __shfl(vrJ[0], src, WAVE), t1 = __shfl(vrJ[1], src, WAVE);
    const float t2 = __shfl(vrJ[2], src, WAVE), t3 = __shfl(vrJ[3], src, WAVE);
    const int r = c & 3;
    return r == 0 ? t0 : r == 1 ? t1 : r == 2 ? t2 : t3;
}

// Sum over the 16 lanes of a row (the column index c); every lane gets it.
__device__ __forceinline__ float row16_sum(float v) {
#pragma unroll
    for (int m = 8; m >= 1; m >>= 1) v += __shfl_xor(v, m, WAVE);
    return v;
}
__device__ __forceinline__ double row16_sum_d(double v) {
#pragma unroll
    for (int m = 8; m >= 1; m >>= 1) v += __shfl_xor(v, m, WAVE);
    return v;
}
// Sum over the 4 row groups (q); every lane gets it.
__device__ __forceinline__ float q4_sum(float v) {
    v += __shfl_xor(v, 16, WAVE);
    return v + __shfl_xor(v, 32, WAVE);
}
__device__ __forceinline__ double q4_sum_d(double v) {
    v += __shfl_xor(v, 16, WAVE);
    return v + __shfl_xor(v, 32, WAVE);
}

// Row f of the augmented tiles (b before the factorisation, z = L^-1 b
// after it), as a column-form vector masked to the first f entries. Row f
// lies in the last block row (f >> 4 == T - 1 for T = ceil((f+1)/16)).
template <int T, int NT>
__device__ __forceinline__ void extract_row_f(const f32x4 (&a)[NT], int f, float (&vc)[T]) {
    const int l = lane(), c = l & 15;
    const int src = (((f & 15) >> 2) << 4) | c, rf = f & 3;
#pragma unroll
    for (int J = 0; J < T; ++J) {
        const float v = __shfl(pick4(a[tile_id(T - 1, J)], rf), src, WAVE);
        vc[J] = (16 * J + c < f) ? v : 0.f;
    }
}

// ------------------------------------------------------------- Cholesky
// Blocked right-looking Cholesky of the leading f x f block of the
// augmented matrix, in place in the accumulator tiles (16-column panels).
//  * Panel: for each pivot column k the raw column is published through
//    LDS (colbuf); the panel's later columns take the rank-1 update
//    A[i][j] -= A[i][k] A[j][k] / A[k][k]. Columns stay unscaled until the
//    panel is done, then every lane scales its own column by 1/sqrt(pivot).
//  * Trailing tiles: A_IJ -= L_IK L_JK^T for the whole panel at once on FP32
//    MFMA (16x16x4), the panel tiles passing through LDS to get the MFMA
//    operand layout.
// Row f of the result is z = L^-1 b. ok becomes false if a pivot was <= 0.
template <int T, int K, int R, int NT>
__device__ __forceinline__ void chol_step(f32x4 (&a)[NT], int cc4, int kend, float *colbuf, bool &ok,
                                          float &myinv) {
    const int cc = cc4 + R;  // cc & 3 == R: static register selection below
    if (cc >= kend) return;
    const int l = lane(), q = l >> 4, c = l & 15;
    const float d = bcast_lane(a[tile_id(K, K)][R], (cc4 << 2) | cc);
    const bool good = d > 0.f;
    ok = ok && good;
    const float rs = good ? __builtin_amdgcn_rsqf(d) : 0.f;
    if (c == cc) {
        myinv = rs;
#pragma unroll
        for (int I = K; I < T; ++I) {
            f32x4 v = a[tile_id(I, K)];
            if (I == K) {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (4 * q + r < cc) v[r] = 0.f;
            }
            *reinterpret_cast<f32x4 *>(colbuf + 16 * I + 4 * q) = v;
        }
    }
    __syncthreads();
    const float lk = (c > cc) ? colbuf[16 * K + c] * (rs * rs) : 0.f;
#pragma unroll
    for (int I = K; I < T; ++I) a[tile_id(I, K)] -= *reinterpret_cast<const f32x4 *>(colbuf + 16 * I + 4 * q) * lk;
    __syncthreads();
}

template <int T, int K, int NT>
__device__ __forceinline__ void chol_panel(f32x4 (&a)[NT], int f, float *colbuf, float *panel, bool &ok) {
    if constexpr (K < T) {
        if (16 * K < f) {
            const int l = lane(), q = l >> 4, c = l & 15;
            const int kend = min(16, f - 16 * K);
            float myinv = 1.f;
            for (int cc4 = 0; cc4 < kend; cc4 += 4) {
                chol_step<T, K, 0>(a, cc4, kend, colbuf, ok, myinv);
                chol_step<T, K, 1>(a, cc4, kend, colbuf, ok, myinv);
                chol_step<T, K, 2>(a, cc4, kend, colbuf, ok, myinv);
                chol_step<T, K, 3>(a, cc4, kend, colbuf, ok, myinv);
            }
#pragma unroll
            for (int I = K; I < T; ++I) a[tile_id(I, K)] *= myinv;
            if constexpr (K + 1 < T) {
                constexpr int TS = 16 * 17;  // one panel tile, rows padded to 17
#pragma unroll
                for (int I = K + 1; I < T; ++I)
#pragma unroll
                    for (int r = 0; r < 4; ++r) panel[(I - K - 1) * TS + (4 * q + r) * 17 + c] = a[tile_id(I, K)][r];
                __syncthreads();
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    float F[T];
#pragma unroll
                    for (int I = K + 1; I < T; ++I) F[I] = panel[(I - K - 1) * TS + (l & 15) * 17 + 4 * s + (l >> 4)];
#pragma unroll
                    for (int I = K + 1; I < T; ++I) {
                        const float nf = -F[I];
#pragma unroll
                        for (int J = K + 1; J <= I; ++J)
                            a[tile_id(I, J)] = __builtin_amdgcn_mfma_f32_16x16x4f32(nf, F[J], a[tile_id(I, J)], 0, 0, 0);
                    }
                }
                __syncthreads();
            }
        }
        chol_panel<T, K + 1>(a, f, colbuf, panel, ok);
    }
}

template <int T, int NT>
__device__ __forceinline__ bool chol_regs(f32x4 (&a)[NT], int f, float *colbuf, float *panel) {
    bool ok = true;
    chol_panel<T, 0>(a, f, colbuf, panel, ok);
    return ok;
}

// -------------------------------------------------- triangular solves
// On the factor held in the tiles (L[16I+4q+r][16J+c] in a[tile(I,J)][r]),
// block by block: the diagonal 16x16 block is solved column by column with
// cross-lane broadcasts, the off-diagonal blocks enter as tile mat-vecs.

// Diagonal block of L^T x = y (column form in/out): column j, j = jend-1..0.
template <int R>
__device__ __forceinline__ void bsub_step(const f32x4 &D, int j4, int jend, float &y, float &xk) {
    const int j = j4 + R;
    if (j >= jend) return;
    const int c = lane() & 15;
    const float yj = bcast_lane(y, j);
    const float ljj = bcast_lane(D[R], (j4 << 2) | j);
    const float xj = yj * __builtin_amdgcn_rcpf(ljj);
    xk = (c == j) ? xj : xk;
    const float lji = __shfl(D[R], (j4 << 2) | c, WAVE);  // L[j][c] of this block
    y = (c < j) ? y - lji * xj : y;
}

// L^T x = z. v: column form, z in, x out (entries >= f are zero). The
// solved blocks also go to the wave's LDS vector buffer xbuf (16T floats),
// where later blocks read them in row form with one 16-byte load; on return
// xbuf holds x.
template <int T, int NT>
__device__ __forceinline__ void back_subst(const f32x4 (&a)[NT], int f, float (&v)[T], float *xbuf) {
    const int l = lane(), q = l >> 4, c = l & 15;
    const int KP = (f - 1) >> 4;
#pragma unroll
    for (int K = T - 1; K >= 0; --K) {
        if (K > KP) {
            v[K] = 0.f;
            if (q == 0) xbuf[16 * K + c] = 0.f;
            __syncthreads();
            continue;
        }
        float p = 0.f;
#pragma unroll
        for (int I = K + 1; I < T; ++I) {
            const f32x4 xr = *reinterpret_cast<const f32x4 *>(xbuf + 16 * I + 4 * q);
#pragma unroll
            for (int r = 0; r < 4; ++r) p += a[tile_id(I, K)][r] * xr[r];
        }
        float y = v[K] - q4_sum(p);
        const int jend = min(16, f - 16 * K);
        float xk = 0.f;
        for (int j4 = (jend - 1) & ~3; j4 >= 0; j4 -= 4) {
            bsub_step<3>(a[tile_id(K, K)], j4, jend, y, xk);
            bsub_step<2>(a[tile_id(K, K)], j4, jend, y, xk);
            bsub_step<1>(a[tile_id(K, K)], j4, jend, y, xk);
            bsub_step<0>(a[tile_id(K, K)], j4, jend, y, xk);
        }
        v[K] = xk;
        if (q == 0) xbuf[16 * K + c] = xk;
        __syncthreads();
    }
}

// Diagonal block of L w = y (row form y, w column form): column j = 0..jend-1.
template <int R>
__device__ __forceinline__ void fsub_step(const f32x4 &D, int j4, int jend, f32x4 &y, float &wk) {
    const int j = j4 + R;
    if (j >= jend) return;
    const int l = lane(), q = l >> 4, c = l & 15;
    const float yj = bcast_lane(y[R], j4 << 2);
    const float ljj = bcast_lane(D[R], (j4 << 2) | j);
    const float wj = yj * __builtin_amdgcn_rcpf(ljj);
    wk = (c == j) ? wj : wk;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float lij = __shfl(D[k], (l & 48) | j, WAVE);  // L[4q+k][j] of this block
        y[k] = (4 * q + k > j) ? y[k] - lij * wj : y[k];
    }
}

// Forward substitution for block K given the block's right-hand side y (row
// form, already reduced by the earlier blocks). Returns w_K (column form).
template <int T, int K, int NT>
__device__ __forceinline__ float fsub_block(const f32x4 (&a)[NT], int f, f32x4 y) {
    const int jend = min(16, f - 16 * K);
    float wk = 0.f;
    for (int j4 = 0; j4 < jend; j4 += 4) {
        fsub_step<0>(a[tile_id(K, K)], j4, jend, y, wk);
        fsub_step<1>(a[tile_id(K, K)], j4, jend, y, wk);
        fsub_step<2>(a[tile_id(K, K)], j4, jend, y, wk);
        fsub_step<3>(a[tile_id(K, K)], j4, jend, y, wk);
    }
    return wk;
}

template <int T, int K, int NT>
__device__ __forceinline__ void fwd_blocks(const f32x4 (&a)[NT], int f, const f32x4 (&vr)[T], float (&w)[T]) {
    if constexpr (K < T) {
        w[K] = 0.f;
        if (16 * K < f) {
            f32x4 u = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int J = 0; J < K; ++J) u += a[tile_id(K, J)] * w[J];
            f32x4 y = vr[K];
#pragma unroll
            for (int r = 0; r < 4; ++r) y[r] -= row16_sum(u[r]);
            w[K] = fsub_block<T, K>(a, f, y);
        }
        fwd_blocks<T, K + 1>(a, f, vr, w);
    }
}

// L w = r. vr: row form in (entries >= f ignored); w: column form out.
template <int T, int NT>
__device__ __forceinline__ void fwd_subst(const f32x4 (&a)[NT], int f, const f32x4 (&vr)[T], float (&w)[T]) {
    fwd_blocks<T, 0>(a, f, vr, w);
}

// ------------------------------------------------------------ mat-vecs
// y = A v for the symmetric matrix held as lower tiles (full diagonal
// tiles) in the accumulator layout, restricted to the leading f x f block.
// v in column form (zero past f); y returned in column form (zero past f).
// Lower tile (I,J), lane (q,c): A[16I+4q+r][16J+c] contributes
//   to y[16I+4q+r] with v[16J+c]   (row partial, summed over c)
//   to y[16J+c]   with v[16I+4q+r] (column partial, summed over q; J < I)
template <int T, int NT>
__device__ __forceinline__ void matvec_f32(const f32x4 (&A)[NT], int f, const float (&v)[T], float (&y)[T]) {
    const int c = lane() & 15, q = lane() >> 4;
    f32x4 vr[T];
#pragma unroll
    for (int I = 0; I < T; ++I) vr[I] = col_to_row(v[I]);
    f32x4 pr[T];
    float pc[T];
#pragma unroll
    for (int I = 0; I < T; ++I) {
        pr[I] = f32x4{0.f, 0.f, 0.f, 0.f};
        pc[I] = 0.f;
    }
#pragma unroll
    for (int I = 0; I < T; ++I)
#pragma unroll
        for (int J = 0; J <= I; ++J) {
            const f32x4 t = A[tile_id(I, J)];
            pr[I] += t * v[J];
            if (J < I) pc[J] += t[0] * vr[I][0] + t[1] * vr[I][1] + t[2] * vr[I][2] + t[3] * vr[I][3];
        }
#pragma unroll
    for (int I = 0; I < T; ++I) {
#pragma unroll
        for (int r = 0; r < 4; ++r) pr[I][r] = row16_sum(pr[I][r]);
        const float yc = row_to_col(pr[I]) + q4_sum(pc[I]);
        y[I] = (16 * I + c < f) ? yc : 0.f;
    }
    (void)q;
}

// r = b - A x with A read back from this wave's global scratch copy of the
// tiles; products and sums in FP64. x, b in column form; r returned in row
// form (FP32, zero past f) for the forward substitution.
template <int T>
__device__ __forceinline__ void residual_fp64(const float *Ag, int f, const float (&b)[T], const float (&x)[T],
                                              f32x4 (&r)[T]) {
    const int l = lane(), q = l >> 4;
    f32x4 xr[T];
#pragma unroll
    for (int I = 0; I < T; ++I) xr[I] = col_to_row(x[I]);
    // One output block J at a time: y_J = sum_{I>=J} A_IJ^T x_I (column
    // partials; the diagonal tile is full and symmetric) + sum_{I<J} A_JI x_I
    // (row partials). Tiles are re-read from the scratch copy (L2-resident).
#pragma unroll
    for (int J = 0; J < T; ++J) {
        double pc = 0.0, pr[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int I = J; I < T; ++I) {
            const f32x4 t = *reinterpret_cast<const f32x4 *>(Ag + tile_id(I, J) * 256 + 4 * l);
#pragma unroll
            for (int k = 0; k < 4; ++k) pc += (double)t[k] * (double)xr[I][k];
        }
#pragma unroll
        for (int I = 0; I < J; ++I) {
            const f32x4 t = *reinterpret_cast<const f32x4 *>(Ag + tile_id(J, I) * 256 + 4 * l);
            const double xi = (double)x[I];
#pragma unroll
            for (int k = 0; k < 4; ++k) pr[k] += (double)t[k] * xi;
        }
        const double pcs = q4_sum_d(pc);        // column form: y_J[c]
        const f32x4 br = col_to_row(b[J]);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double colpart = __shfl(pcs, (l & 48) | (4 * q + k), WAVE);  // -> row form
            const int i = 16 * J + 4 * q + k;
            const double y = row16_sum_d(pr[k]) + colpart;
            r[J][k] = i < f ? (float)((double)br[k] - y) : 0.f;
        }
    }
}

// One step of iterative refinement, x += A^-1 (b - A x), with the residual
// in FP64 and the correction solved on the FP32 factor. The residual of
// block K is formed just before block K of the forward substitution needs
// it, so only one block of it is ever live. Ag is this wave's scratch copy
// of the augmented tiles before the factorisation (row f holds b).
template <int T, int K, int NT>
__device__ __forceinline__ void refine_blocks(const f32x4 (&a)[NT], int f, const float *Ag, const float (&x)[T],
                                              const float *xbuf, float (&w)[T]) {
    if constexpr (K < T) {
        w[K] = 0.f;
        // Keep the scratch-tile loads of this block from being hoisted into
        // earlier blocks (that would hold every tile of A in registers).
        __builtin_amdgcn_sched_barrier(0);
        if (16 * K < f) {
            const int l = lane(), q = l >> 4;
            const int bsrc = ((f & 15) >> 2) << 4, brf = f & 3;
            double pc = 0.0, pr[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int I = K; I < T; ++I) {
                const f32x4 t = *reinterpret_cast<const f32x4 *>(Ag + tile_id(I, K) * 256 + 4 * l);
                const f32x4 xr = *reinterpret_cast<const f32x4 *>(xbuf + 16 * I + 4 * q);
#pragma unroll
                for (int k = 0; k < 4; ++k) pc += (double)t[k] * (double)xr[k];
            }
#pragma unroll
            for (int I = 0; I < K; ++I) {
                const f32x4 t = *reinterpret_cast<const f32x4 *>(Ag + tile_id(K, I) * 256 + 4 * l);
                const double xi = (double)x[I];
#pragma unroll
                for (int k = 0; k < 4; ++k) pr[k] += (double)t[k] * xi;
            }
            const double pcs = q4_sum_d(pc);
            const float *brow = Ag + tile_id(T - 1, K) * 256;  // row f lies in block T-1
            f32x4 u = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int J = 0; J < K; ++J) u += a[tile_id(K, J)] * w[J];
            f32x4 y;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int i = 16 * K + 4 * q + k;
                const double colpart = __shfl(pcs, (l & 48) | (4 * q + k), WAVE);
                const double bk = (double)brow[4 * (bsrc | (4 * q + k)) + brf];
                const double rs = row16_sum_d(pr[k]);
                y[k] = (i < f ? (float)(bk - rs - colpart) : 0.f) - row16_sum(u[k]);
            }
            w[K] = fsub_block<T, K>(a, f, y);
        }
        refine_blocks<T, K + 1>(a, f, Ag, x, xbuf, w);
    }
}

template <int T, int NT>
__device__ __forceinline__ void refine_step(const f32x4 (&a)[NT], int f, const float *Ag, float (&x)[T],
                                            float *xbuf) {
    float w[T];
    refine_blocks<T, 0>(a, f, Ag, x, xbuf, w);
    back_subst<T>(a, f, w, xbuf);  // xbuf now holds the correction
    const int q = lane() >> 4, c = lane() & 15;
#pragma unroll
    for (int J = 0; J < T; ++J) {
        x[J] += w[J];
        if (q == 0) xbuf[16 * J + c] = x[J];
    }
    __syncthreads();
}

__device__ __forceinline__ float colform_dot(const float *a, const float *b, int T) {
    float s = 0.f;
    for (int J = 0; J < T; ++J) s += a[J] * b[J];
    return row16_sum(s);
}

// cuMF CG (cumf_als_cg/cg.cu:37-232) on the register tiles: warm start x,
// `iters` iterations, exit when r.r < tol after an update; p.Ap == 0 stops.
template <int T, int NT>
__device__ __forceinline__ void cg_regs(const f32x4 (&A)[NT], int f, const float (&b)[T], float (&x)[T], int iters,
                                        float tol) {
    float r[T], p[T], ap[T];
    matvec_f32<T>(A, f, x, ap);
#pragma unroll
    for (int J = 0; J < T; ++J) {
        r[J] = b[J] - ap[J];
        p[J] = r[J];
    }
    float rsold = colform_dot(r, r, T);
    for (int it = 0; it < iters; ++it) {
        matvec_f32<T>(A, f, p, ap);
        const float pap = colform_dot(p, ap, T);
        if (pap == 0.f) break;
        const float alpha = rsold / pap;
#pragma unroll
        for (int J = 0; J < T; ++J) {
            x[J] += alpha * p[J];
            r[J] -= alpha * ap[J];
        }
        const float rsnew = colform_dot(r, r, T);
        if (rsnew < tol) break;
        const float beta = rsnew / rsold;
        rsold = rsnew;
#pragma unroll
        for (int J = 0; J < T; ++J) p[J] = r[J] + beta * p[J];
    }
}

// Column-form vector -> one row of a 16T-strided matrix in global memory
// (row-group q stores blocks q, q+4, ...; 16 consecutive floats each).
template <int T>
__device__ __forceinline__ void store_colform(float *dst, const float (&v)[T]) {
    const int l = lane(), q = l >> 4, c = l & 15;
#pragma unroll
    for (int J4 = 0; J4 < T; J4 += 4) {
        const int J = J4 + q;
        float val = 0.f;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (J4 + k < T && q == k) val = v[J4 + k];
        if (J < T) dst[16 * J + c] = val;
    }
}
template <int T, int E>
__device__ __forceinline__ void store_colform16(uint16_t *dst, const float (&v)[T]) {
    const int l = lane(), q = l >> 4, c = l & 15;
#pragma unroll
    for (int J4 = 0; J4 < T; J4 += 4) {
        const int J = J4 + q;
        float val = 0.f;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (J4 + k < T && q == k) val = v[J4 + k];
        if (J < T) dst[16 * J + c] = to16<E>(val);
    }
}
template <int T>
__device__ __forceinline__ void load_colform(const float *src, int f, float (&v)[T]) {
    const int c = lane() & 15;
#pragma unroll
    for (int J = 0; J < T; ++J) v[J] = (16 * J + c < f) ? src[16 * J + c] : 0.f;
}

// ------------------------------------------------------------- the kernel
template <int T, int E, int S>
__device__ __forceinline__ void update_body(const UpdateArgs &a) {
    using G = Geo<T, E>;
    extern __shared__ __attribute__((aligned(16))) char lds[];
    float *colbuf = reinterpret_cast<float *>(lds) + 64;
    float *xbuf = colbuf + 128;
    float *stage = reinterpret_cast<float *>(lds + G::MISC_FLOATS * 4);  // reused as the panel buffer
    const int l = lane();
    const int f = a.f;
    float *Ascr = a.scratch ? a.scratch + (int64_t)blockIdx.x * G::NT * 256 : nullptr;
    f32x4 acc[G::NT];

    for (;;) {
        int b = 0;
        if (l == 0) b = atomicAdd(a.queue, 1);
        b = __builtin_amdgcn_readfirstlane(__shfl(b, 0, WAVE));
        if (b >= a.nbatches) break;
        const int u0 = a.batch_begin[b], u1 = a.batch_begin[b + 1];
        for (int u = u0; u < u1; ++u) {
            const int4 w = a.units[u];
            const int row = w.x;
            const int n = a.ptr[row + 1] - a.ptr[row];  // whole row (lambda * n_u)
            float x[T];
#pragma unroll
            for (int J = 0; J < T; ++J) x[J] = 0.f;
            if (n > 0) {
                gram_row<T, E>(a, w.y, w.z - w.y, acc, lds);
                if (w.w >= 0) {
                    // Split row: publish this part's tiles; the last part to
                    // arrive sums every part in slot order and solves.
                    const int4 sp = a.splits[w.w];
                    const int part = (w.y - a.ptr[row]) / sp.z;
                    float *mine = a.partial + (int64_t)(sp.x + part) * G::NT * 256;
#pragma unroll
                    for (int t = 0; t < G::NT; ++t) *reinterpret_cast<f32x4 *>(mine + t * 256 + 4 * l) = acc[t];
                    __threadfence();
                    int ticket = 0;
                    if (l == 0) ticket = atomicAdd(a.counters + w.w, 1);
                    ticket = __builtin_amdgcn_readfirstlane(__shfl(ticket, 0, WAVE));
                    if (ticket != sp.y - 1) continue;
                    __threadfence();
#pragma unroll
                    for (int t = 0; t < G::NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
                    for (int p = 0; p < sp.y; ++p) {
                        const float *src = a.partial + (int64_t)(sp.x + p) * G::NT * 256;
#pragma unroll
                        for (int t = 0; t < G::NT; ++t)
                            acc[t] += *reinterpret_cast<const f32x4 *>(src + t * 256 + 4 * l);
                    }
                    if (l == 0) atomicExch(a.counters + w.w, 0);
                }
                const float lam = a.lambda_rows ? a.lambda_rows[row] : a.lambda;
                add_diag<T>(acc, f, lam * (float)n);
                bool ok = true;
                if constexpr (S == SOLVE_CHOL || S == SOLVE_CHOL_REFINE) {
                    if constexpr (S == SOLVE_CHOL_REFINE) {
#pragma unroll
                        for (int t = 0; t < G::NT; ++t) *reinterpret_cast<f32x4 *>(Ascr + t * 256 + 4 * l) = acc[t];
                    }
                    ok = chol_regs<T>(acc, f, colbuf, stage);
                    extract_row_f<T>(acc, f, x);
                    back_subst<T>(acc, f, x, xbuf);
                    if constexpr (S == SOLVE_CHOL_REFINE) {
                        for (int s = 0; s < a.refine; ++s) refine_step<T>(acc, f, Ascr, x, xbuf);
                    }
                    if (a.Lsave) {
                        float *dst = a.Lsave + (int64_t)row * G::NT * 256;
#pragma unroll
                        for (int t = 0; t < G::NT; ++t) *reinterpret_cast<f32x4 *>(dst + t * 256 + 4 * l) = acc[t];
                    }
                } else {
                    float bvec[T];
                    extract_row_f<T>(acc, f, bvec);
                    if (a.cg_fp16_a) {
#pragma unroll
                        for (int t = 0; t < G::NT; ++t)
#pragma unroll
                            for (int r = 0; r < 4; ++r) acc[t][r] = (float)(_Float16)acc[t][r];
                    }
                    load_colform<T>(a.X + (int64_t)row * G::FP, f, x);
                    cg_regs<T>(acc, f, bvec, x, a.cg_iters, a.cg_tol);
                }
                bool finite = true;
#pragma unroll
                for (int J = 0; J < T; ++J) finite = finite && isfinite(x[J]);
                const bool bad = !ok || __any(!finite);
                if (bad) {
                    if (l == 0) atomicAdd(a.fail_count, 1);
#pragma unroll
                    for (int J = 0; J < T; ++J) x[J] = 0.f;
                }
            }
            store_colform<T>(a.X + (int64_t)row * G::FP, x);
            if constexpr (E != PREC_FP32) {
                if (a.Xs) store_colform16<T, E>(a.Xs + (int64_t)row * G::FP, x);
            }
        }
    }
    // Exit protocol: the last wave out resets the queue for the next launch.
    if (l == 0) {
        __threadfence();
        const int done = atomicAdd(a.queue + 1, 1);
        if (done == (int)gridDim.x - 1) {
            atomicExch(a.queue, 0);
            atomicExch(a.queue + 1, 0);
        }
    }
}

template <int T, int E, int S>
__global__ __launch_bounds__(WAVE) void update_kernel(UpdateArgs a) {
    update_body<T, E, S>(a);
}

// The refinement variant holds the factor, the solution and the FP64
// residual pieces at once; at two waves per SIMD it trades a small spill
// for twice the latency hiding of the one-wave form.
template <int T, int E, int S>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(2, 2))) void update_kernel_w2(UpdateArgs a) {
    update_body<T, E, S>(a);
}

}  // namespace als
