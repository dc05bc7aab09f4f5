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
    static constexpr int NP = FP * (FP + 1) / 2;// packed lower triangle
    static constexpr int YT_H = 40;             // 16-bit stage row: 32 ratings + 8 pad (80 B)
    static constexpr int YR = (T & 1) ? FP : FP + 16;  // FP32 stage row stride (floats)
    static constexpr int STAGE_BYTES = E == PREC_FP32 ? CHUNK * YR * 4 : FP * YT_H * 2;
    static constexpr int MISC_FLOATS = 32 + 32 + FP + 128;  // ibuf, rbuf, colbuf, vbuf
    static constexpr int R1_BYTES = (STAGE_BYTES > NP * 4 ? STAGE_BYTES : NP * 4);
    __host__ __device__ static constexpr int lds_bytes(bool second_packed) {
        return MISC_FLOATS * 4 + R1_BYTES + (second_packed ? NP * 4 : 0);
    }
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
                int id = ibuf[k];
                f32x4 v = {0.f, 0.f, 0.f, 0.f};
                if (id >= 0) v = *reinterpret_cast<const f32x4 *>(a.Y + (int64_t)id * G::FP + 4 * s);
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
                    int id = ibuf[8 * g + e];
                    u32x4 z = {0u, 0u, 0u, 0u};
                    if (id >= 0) z = *reinterpret_cast<const u32x4 *>(Ysh + (int64_t)id * G::FP + 8 * s);
                    d[e] = z;
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    u32x4 lo, hi;
#pragma unroll
                    for (int w = 0; w < 4; ++w) {
                        uint32_t x0 = d[2 * w][v], x1 = d[2 * w + 1][v];
                        lo[w] = (x0 & 0xffffu) | (x1 << 16);
                        hi[w] = (x0 >> 16) | (x1 & 0xffff0000u);
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
        for (int r = 0; r < 4; ++r)
            if (4 * q + r == c) {
                int d = 16 * I + c;
                if (d < f) acc[tile_id(I, I)][r] += reg;
                else if (d > f) acc[tile_id(I, I)][r] = 1.f;
            }
}

// Lower triangle of the register tiles -> packed row-major LDS
// (P[i(i+1)/2 + j], j <= i). With ROUND16 the values pass through FP16
// storage (the "FP16-stored A" of the reference's CG_FP16 solver).
template <int T, int NT, bool ROUND16 = false>
__device__ __forceinline__ void dump_packed(const f32x4 (&acc)[NT], float *P, int f = 0) {
    const int l = lane(), q = l >> 4, c = l & 15;
#pragma unroll
    for (int I = 0; I < T; ++I)
#pragma unroll
        for (int J = 0; J <= I; ++J)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 16 * I + 4 * q + r, j = 16 * J + c;
                if (J < I || c <= 4 * q + r) {
                    float v = acc[tile_id(I, J)][r];
                    if (ROUND16 && i != f) v = (float)(_Float16)v;  // A in FP16, b stays FP32
                    P[i * (i + 1) / 2 + j] = v;
                }
            }
}

// ------------------------------------------------------------- Cholesky
// Right-looking Cholesky of the leading f x f block of the augmented matrix,
// in place in the accumulator tiles. Column k is published through LDS
// (colbuf) once per step; every lane then applies the rank-1 update to its
// 4 rows x 1 column of every trailing tile. Row f of the result is
// z = L^-1 b. Returns false if a pivot was not positive.
template <int T, int K, int NT>
__device__ __forceinline__ void chol_block(f32x4 (&a)[NT], int f, float *colbuf, bool &ok) {
    if constexpr (K < T) {
        const int l = lane(), q = l >> 4, c = l & 15;
        const int kend = min(16, f - 16 * K);
        for (int cc = 0; cc < kend; ++cc) {
            const float d = bcast_lane(pick4(a[tile_id(K, K)], cc & 3), (cc >> 2) * 16 + cc);
            const bool good = d > 0.f;
            ok = ok && good;
            const float inv = good ? 1.0f / sqrtf(d) : 0.f;
            if (c == cc) {
#pragma unroll
                for (int I = K; I < T; ++I) {
                    a[tile_id(I, K)] *= inv;
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
            f32x4 lrow[T - K];
            float lcol[T - K];
#pragma unroll
            for (int I = K; I < T; ++I) {
                lrow[I - K] = *reinterpret_cast<const f32x4 *>(colbuf + 16 * I + 4 * q);
                lcol[I - K] = colbuf[16 * I + c];
            }
            if (c <= cc) lcol[0] = 0.f;
#pragma unroll
            for (int I = K; I < T; ++I)
#pragma unroll
                for (int J = K; J <= I; ++J) a[tile_id(I, J)] -= lrow[I - K] * lcol[J - K];
            __syncthreads();
        }
        chol_block<T, K + 1>(a, f, colbuf, ok);
    }
}

template <int T, int NT>
__device__ __forceinline__ bool chol_regs(f32x4 (&a)[NT], int f, float *colbuf) {
    bool ok = true;
    chol_block<T, 0>(a, f, colbuf, ok);
    return ok;
}

// Triangular solves on a packed lower factor P (row-major, P[i(i+1)/2+j]).
// Vectors live one element per lane: v_lo = v[lane], v_hi = v[64 + lane].

// L^T x = z, in place (z in, x out).
__device__ __forceinline__ void solve_lt(const float *P, int f, float &vlo, float &vhi) {
    const int l = lane();
    for (int k = f - 1; k >= 0; --k) {
        const int kk = k * (k + 1) / 2;
        const float zk = k < WAVE ? bcast_lane(vlo, k) : bcast_lane(vhi, k - WAVE);
        const float xk = zk * __builtin_amdgcn_rcpf(P[kk + k]);
        vlo = (l == k) ? xk : vlo;
        vhi = (WAVE + l == k) ? xk : vhi;
        if (l < k) vlo -= P[kk + l] * xk;
        if (WAVE + l < k) vhi -= P[kk + WAVE + l] * xk;
    }
}

// L w = r, in place (r in, w out).
__device__ __forceinline__ void solve_l(const float *P, int f, float &vlo, float &vhi) {
    const int l = lane();
    const int il = l, ih = WAVE + l;
    const int bl = il * (il + 1) / 2, bh = ih * (ih + 1) / 2;
    for (int k = 0; k < f; ++k) {
        const int kk = k * (k + 1) / 2;
        const float rk = k < WAVE ? bcast_lane(vlo, k) : bcast_lane(vhi, k - WAVE);
        const float wk = rk * __builtin_amdgcn_rcpf(P[kk + k]);
        vlo = (l == k) ? wk : vlo;
        vhi = (WAVE + l == k) ? wk : vhi;
        if (il > k && il < f) vlo -= P[bl + k] * wk;
        if (ih > k && ih < f) vhi -= P[bh + k] * wk;
    }
}

// y = A p for a packed symmetric A (lane-per-row); p passes through vbuf.
__device__ __forceinline__ void matvec_packed(const float *A, int f, float plo, float phi, float *vbuf,
                                              float &ylo, float &yhi) {
    const int l = lane();
    vbuf[l] = plo;
    vbuf[WAVE + l] = phi;
    __syncthreads();
    const int il = l, ih = WAVE + l;
    const int bl = il * (il + 1) / 2, bh = ih * (ih + 1) / 2;
    float sl = 0.f, sh = 0.f;
    for (int j = 0; j < f; ++j) {
        const float pj = vbuf[j];
        const int bj = j * (j + 1) / 2;
        if (il < f) sl += A[j <= il ? bl + j : bj + il] * pj;
        if (ih < f) sh += A[j <= ih ? bh + j : bj + ih] * pj;
    }
    __syncthreads();
    ylo = sl;
    yhi = sh;
}

// r = b - A x in FP64 (b is row f of the packed augmented matrix).
__device__ __forceinline__ void residual_fp64(const float *A, int f, float xlo, float xhi, float *vbuf,
                                              double &rlo, double &rhi) {
    const int l = lane();
    vbuf[l] = xlo;
    vbuf[WAVE + l] = xhi;
    __syncthreads();
    const int il = l, ih = WAVE + l;
    const int bl = il * (il + 1) / 2, bh = ih * (ih + 1) / 2, bf = f * (f + 1) / 2;
    double sl = il < f ? (double)A[bf + il] : 0.0, sh = ih < f ? (double)A[bf + ih] : 0.0;
    for (int j = 0; j < f; ++j) {
        const double xj = (double)vbuf[j];
        const int bj = j * (j + 1) / 2;
        if (il < f) sl -= (double)A[j <= il ? bl + j : bj + il] * xj;
        if (ih < f) sh -= (double)A[j <= ih ? bh + j : bj + ih] * xj;
    }
    __syncthreads();
    rlo = sl;
    rhi = sh;
}

// cuMF CG (cumf_als_cg/cg.cu:37-232): warm start from x, cg_iters
// iterations, exit when r.r < tol after an update; p.Ap == 0 also stops.
__device__ __forceinline__ void cg_packed(const float *A, int f, int iters, float tol, float *vbuf,
                                          float &xlo, float &xhi) {
    const int l = lane(), bf = f * (f + 1) / 2;
    const float blo = l < f ? A[bf + l] : 0.f, bhi = WAVE + l < f ? A[bf + WAVE + l] : 0.f;
    float axl, axh;
    matvec_packed(A, f, xlo, xhi, vbuf, axl, axh);
    float rlo = blo - axl, rhi = bhi - axh;
    float plo = rlo, phi = rhi;
    float rsold = wave_sum(rlo * rlo + rhi * rhi);
    for (int it = 0; it < iters; ++it) {
        float aplo, aphi;
        matvec_packed(A, f, plo, phi, vbuf, aplo, aphi);
        const float pap = wave_sum(plo * aplo + phi * aphi);
        if (pap == 0.f) break;
        const float alpha = rsold / pap;
        xlo += alpha * plo;
        xhi += alpha * phi;
        rlo -= alpha * aplo;
        rhi -= alpha * aphi;
        const float rsnew = wave_sum(rlo * rlo + rhi * rhi);
        if (rsnew < tol) break;
        const float beta = rsnew / rsold;
        rsold = rsnew;
        plo = rlo + beta * plo;
        phi = rhi + beta * phi;
    }
}

// ------------------------------------------------------------- the kernel
template <int T, int E, int S>
__global__ __launch_bounds__(WAVE) void update_kernel(UpdateArgs a) {
    using G = Geo<T, E>;
    extern __shared__ __attribute__((aligned(16))) char lds[];
    float *colbuf = reinterpret_cast<float *>(lds) + 64;
    float *vbuf = colbuf + G::FP;
    float *R1 = reinterpret_cast<float *>(lds + G::MISC_FLOATS * 4);
    float *R2 = reinterpret_cast<float *>(lds + G::MISC_FLOATS * 4 + G::R1_BYTES);
    const int l = lane();
    const int f = a.f;
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
            float xlo = 0.f, xhi = 0.f;
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
                if constexpr (S == SOLVE_CHOL) {
                    if (a.refine > 0) dump_packed<T>(acc, R2);
                    ok = chol_regs<T>(acc, f, colbuf);
                    dump_packed<T>(acc, R1);
                    __syncthreads();
                    const int bf = f * (f + 1) / 2;
                    xlo = l < f ? R1[bf + l] : 0.f;
                    xhi = WAVE + l < f ? R1[bf + WAVE + l] : 0.f;
                    solve_lt(R1, f, xlo, xhi);
                    for (int s = 0; s < a.refine; ++s) {
                        double rl, rh;
                        residual_fp64(R2, f, xlo, xhi, vbuf, rl, rh);
                        float dl = (float)rl, dh = (float)rh;
                        solve_l(R1, f, dl, dh);
                        solve_lt(R1, f, dl, dh);
                        xlo += dl;
                        xhi += dh;
                    }
                    if (a.Lsave) {
                        float *dst = a.Lsave + (int64_t)row * G::NP;
                        for (int e = l; e < G::NP; e += WAVE) dst[e] = R1[e];
                    }
                } else {
                    if (a.cg_fp16_a) dump_packed<T, G::NT, true>(acc, R1, f);
                    else dump_packed<T, G::NT, false>(acc, R1);
                    __syncthreads();
                    xlo = l < f ? a.X[(int64_t)row * G::FP + l] : 0.f;
                    xhi = WAVE + l < f ? a.X[(int64_t)row * G::FP + WAVE + l] : 0.f;
                    cg_packed(R1, f, a.cg_iters, a.cg_tol, vbuf, xlo, xhi);
                }
                const bool bad = !ok || __any(!(isfinite(xlo) && isfinite(xhi)));
                if (bad) {
                    if (l == 0) atomicAdd(a.fail_count, 1);
                    xlo = 0.f;
                    xhi = 0.f;
                }
                __syncthreads();
            }
            float *xr = a.X + (int64_t)row * G::FP;
            if (l < G::FP) xr[l] = l < f ? xlo : 0.f;
            if (WAVE + l < G::FP) xr[WAVE + l] = WAVE + l < f ? xhi : 0.f;
            if constexpr (E != PREC_FP32) {
                if (a.Xs) {
                    uint16_t *xs = a.Xs + (int64_t)row * G::FP;
                    if (l < G::FP) xs[l] = l < f ? to16<E>(xlo) : (uint16_t)0;
                    if (WAVE + l < G::FP) xs[WAVE + l] = WAVE + l < f ? to16<E>(xhi) : (uint16_t)0;
                }
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

}  // namespace als
