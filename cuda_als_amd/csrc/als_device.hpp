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

#include <type_traits>

#include "als_launch.hpp"

namespace als {

constexpr int WAVE = 64;
constexpr int CHUNK = 32;

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

__host__ __device__ constexpr int tile_id(int I, int J) { return I * (I + 1) / 2 + J; }

__device__ __forceinline__ int lane() { return (int)threadIdx.x; }

// The lane index recomputed where it is used (two VALU instructions): in a
// loop that holds most registers, the compiler would otherwise keep
// threadIdx.x in a spill slot and reload it from memory every iteration.
__device__ __forceinline__ int lane_here() {
    int v;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(v));
    return v;
}

__device__ __forceinline__ float bcast_lane(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

// LDS hand-off between the lanes of one wave. Every workgroup of the update
// kernels is exactly one wave, and LDS executes a wave's operations in
// issue order, so a later read sees an earlier write of any lane once the
// compiler keeps them in order. No s_barrier, and no wait for outstanding
// global loads (which __syncthreads' workgroup fence would force, draining
// the gathers prefetched for the next chunk).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
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
#ifndef ALS_DUAL_TD_MAX
#define ALS_DUAL_TD_MAX 5
#endif

template <int T, int E> struct Geo {
    static constexpr int NT = T * (T + 1) / 2;  // lower-triangle 16x16 tiles
    static constexpr int FP = 16 * T;           // padded augmented dimension
    static constexpr int YT_H = 40;             // 16-bit stage row: 32 ratings + 8 pad (80 B)
    // Factor row j of the 16-bit stage sits at row j + j / 8: the gather
    // writes rows 8s + 2v (s = 0..12 in one instruction) and the MFMA reads
    // rows 16I + c, and the skew spreads both over all eight 16-byte bank
    // groups of a 128-byte LDS line (without it the writes of one
    // instruction share four).
    __host__ __device__ static constexpr int yt_row(int j) { return j + (j >> 3); }
    static constexpr int YR = (T & 1) ? FP : FP + 16;  // FP32 stage row stride (floats)
    // Rating-major 16-bit stage of the pipelined Gram (row k = rating k of
    // the chunk, column j = factor j): row stride in halves. A transposing
    // operand read takes 32 bytes from each of four consecutive rows; with
    // 2 * RS16 bytes an odd multiple of 32 modulo 256, or T = 2 mod 4, the
    // four land on distinct LDS banks.
    static constexpr int RS16 = (T % 4 == 0) ? FP + 16 : FP;
    // Rows with fewer than 16 * DUAL_TD ratings take the dual system (T >= 5).
    static constexpr int DUAL_TD = T >= 5 ? (T - 2 < ALS_DUAL_TD_MAX ? T - 2 : ALS_DUAL_TD_MAX) : 0;
    // The double-buffered LDS-DMA Gram (gram_row16_dma): two rating-major
    // buffers of CHUNK x FP halves, then a second ibuf/rbuf pair (256 bytes).
    // Used where a row of FP halves keeps a transposing read's four rows on
    // distinct LDS banks (T % 4 != 0) and a chunk's tasks fit a wave (T <= 8).
    static constexpr bool DMA_GRAM = gram_dma_built(T, E);
    static constexpr int DMA_STAGE_BYTES = 2 * CHUNK * FP * 2 + 256;
    static constexpr int GRAM_STAGE_BYTES =
        E == PREC_FP32 ? CHUNK * YR * 4
                       : (DMA_GRAM && DMA_STAGE_BYTES > yt_row(FP) * YT_H * 2 ? DMA_STAGE_BYTES
                                                                               : yt_row(FP) * YT_H * 2);
    // LDS of the update kernels (misc + Gram stage) and of the dual kernel (misc)
    // ibuf (32 ints), rbuf (32 floats), xs (FP floats), eb (32 doubles)
    static constexpr int MISC_FLOATS = 32 + 32 + FP + 64;
    __host__ __device__ static constexpr int lds_bytes() { return MISC_FLOATS * 4 + GRAM_STAGE_BYTES; }
    // the classic-refinement variant restages FP32 rows into the same stage
    static constexpr int FP32_STAGE_BYTES = CHUNK * YR * 4;
    __host__ __device__ static constexpr int refine_lds_bytes() {
        return MISC_FLOATS * 4 + (GRAM_STAGE_BYTES > FP32_STAGE_BYTES ? GRAM_STAGE_BYTES : FP32_STAGE_BYTES);
    }
    __host__ __device__ static constexpr int dual_lds_bytes() { return MISC_FLOATS * 4; }
};

// ------------------------------------------------------------------ Gram
// Zero the stage columns no chunk writes (padding factors past the loaded
// segments), once per row.
template <int T, int E>
__device__ __forceinline__ void stage_clear(int f, char *stage) {
    using G = Geo<T, E>;
    const int l = lane();
    if constexpr (E == PREC_FP32) {
        float *Ys = reinterpret_cast<float *>(stage);
        const int nseg = (f + 3) >> 2;  // float4 segments carrying data
        for (int e = l; e < CHUNK * (G::FP - 4 * nseg); e += WAVE) {
            int k = e / (G::FP - 4 * nseg), col = 4 * nseg + e % (G::FP - 4 * nseg);
            Ys[k * G::YR + col] = 0.f;
        }
    } else {
        uint16_t *Yt = reinterpret_cast<uint16_t *>(stage);
        const int nseg = (f + 7) >> 3;  // 16-byte segments carrying data
        for (int e = l; e < (G::FP - 8 * nseg) * CHUNK; e += WAVE) {
            int j = 8 * nseg + e / CHUNK, k = e % CHUNK;
            Yt[G::yt_row(j) * G::YT_H + k] = 0;
        }
    }
}

// Stage ratings k0 .. k0+nv-1 of the row starting at rs (nv <= 32): their
// item indices and values to ibuf/rbuf, the gathered rows y_i of the fixed
// factors to the stage (FP32: Ys[rating][factor]; 16-bit: Yt[factor][rating]),
// with the rating in column f. Slots past nv hold zero rows. Ends with the
// stage visible to the whole wave.
template <int T, int E>
__device__ __forceinline__ void stage_chunk(const UpdateArgs &a, int64_t rs, int k0, int nv, char *lds) {
    using G = Geo<T, E>;
    const int l = lane();
    const int f = a.f;
    int *ibuf = reinterpret_cast<int *>(lds);
    float *rbuf = reinterpret_cast<float *>(lds) + 32;
    char *stage = lds + G::MISC_FLOATS * 4;
    if (l < CHUNK) {
        ibuf[l] = l < nv ? a.idx[rs + k0 + l] : -1;
        rbuf[l] = l < nv ? a.val[rs + k0 + l] : 0.f;
    }
    wave_sync();
    if constexpr (E == PREC_FP32) {
        float *Ys = reinterpret_cast<float *>(stage);
        const int nseg = (f + 3) >> 2;
        const int dk = WAVE / nseg, ds = WAVE % nseg;
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
        wave_sync();
        if (l < CHUNK) Ys[l * G::YR + f] = rbuf[l];
    } else {
        uint16_t *Yt = reinterpret_cast<uint16_t *>(stage);
        const uint16_t *Ysh = a.Ys;
        const int nseg = (f + 7) >> 3;
        // Task = (group of 8 ratings g, segment s of 8 factors): eight
        // 16-byte loads, an 8x8 transpose of 16-bit values in registers,
        // eight 16-byte LDS writes of Yt[8s+e][8g..8g+7].
        for (int task = l; task < 4 * nseg; task += WAVE) {
            const int g = task / nseg, s = task - g * nseg;
            u32x4 d[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                // Unused slots of the last chunk read a row of zeros: the
                // load itself is unconditional.
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
                *reinterpret_cast<u32x4 *>(Yt + G::yt_row(8 * s + 2 * v) * G::YT_H + 8 * g) = lo;
                *reinterpret_cast<u32x4 *>(Yt + G::yt_row(8 * s + 2 * v + 1) * G::YT_H + 8 * g) = hi;
            }
        }
        wave_sync();
        if (l < CHUNK) Yt[G::yt_row(f) * G::YT_H + l] = to16<E>(rbuf[l]);
    }
    wave_sync();
}

// Accumulates the lower tiles of the augmented Gram of one row into acc
// (q = lane/16, c = lane%16): diagonal tiles in the MFMA accumulator layout,
// acc[tile_id(I,I)][r] = G[16I + 4q + r][16I + c]; off-diagonal tiles as
// the accumulator of the transposed product, acc[tile_id(I,J)][r] =
// G[16J + 4q + r][16I + c] = G[16I + c][16J + 4q + r] (I > J), the layout
// the register Cholesky works in.
// 16-bit Gram of one row when one gather task per lane covers a chunk
// (4 * nseg <= 64, f <= 128), software-pipelined: the gathers of chunk k+1
// are issued before the MFMAs of chunk k and land in registers while they
// run, and the indices and ratings of chunk k+2 are loaded meanwhile.
// Task (g, s) of a lane: segment s (factors 8s .. 8s+7) of ratings 8g ..
// 8g+7, eight 16-byte loads written as they arrive to a rating-major stage
// Yr[rating][factor]. The MFMA operands come out of it transposed: for
// column block I, lane (q, c) needs A[c][k] = Yr[k][16I + c] for k = 8q ..
// 8q+7, and two ds_read_b64_tr_b16 give it: within each 16-lane group,
// lane i supplies the address of 4 halves, Yr[8q + 4h + i/4][16I + 4(i%4)
// .. +3], and lane j receives element j + 16t of the group's 64, which is
// Yr[8q + 4h + t][16I + j], t = 0..3.
template <int T, int E>
__device__ __forceinline__ void gram_row16_piped(const UpdateArgs &a, int64_t rs, int n,
                                                 f32x4 (&acc)[Geo<T, E>::NT], char *lds) {
    using G = Geo<T, E>;
    constexpr int RS = G::RS16;
    typedef short v4i16 __attribute__((ext_vector_type(4)));
    typedef __attribute__((address_space(3))) v4i16 lds_v4i16;
    const int l = lane();
    const int f = a.f;
    int *ibuf = reinterpret_cast<int *>(lds);
    float *rbuf = reinterpret_cast<float *>(lds) + 32;
    uint16_t *Yr = reinterpret_cast<uint16_t *>(lds + G::MISC_FLOATS * 4);
    const int nseg = (f + 7) >> 3;
    const bool active = l < 4 * nseg;  // task (g, s) of lane l
    const int g = l / nseg, s = l - g * nseg;
    uint16_t *wbase = Yr + 8 * g * RS + 8 * s;
    const uint16_t *rbase = Yr + (8 * (l >> 4) + ((l & 15) >> 2)) * RS + 4 * (l & 3);
    const int64_t soff = 8 * s;
    const int soff2 = 16 * s;  // the same in bytes
    // the shadow as a buffer (raw, byte offsets, zeros past ys_bytes)
    const __amdgpu_buffer_rsrc_t ysr =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t *>(a.Ys), (short)0, a.ys_bytes, 0x00020000);
    if (n <= 0) return;
    // columns no gather writes (8 nseg .. 16T): zero, but for the rating
    // column f, which each chunk writes (when f % 8 != 0 it lies in the last
    // gathered segment, whose padding halves are zero in the shadow rows)
    // (in 16-byte pieces: piece e is row e % 32, columns 8 nseg + 8 (e / 32))
    for (int e = l; e < 4 * (G::FP - 8 * nseg); e += WAVE)
        *reinterpret_cast<u32x4 *>(Yr + (e & 31) * RS + 8 * nseg + 8 * (e >> 5)) = u32x4{0u, 0u, 0u, 0u};
    // Raw indices and ratings of the chunk at k0, loaded by every lane (lane
    // l and l + 32 alike) at positions clamped into the row, so the loads
    // are unconditional; the positions past the row's end are masked when
    // the chunk is published.
    auto load_ids = [&](int k0, int &id_raw, float &v_raw) {
        const int64_t p = rs + min(k0 + (lane_here() & 31), n - 1);
        id_raw = a.idx[p];
        v_raw = a.val[p];
    };
    // Publishes the chunk at k0 (ibuf, rbuf: both halves of the wave write
    // the same values) and gathers its eight rows per active lane into d
    // (slots past the row's end read the zero row).
    u32x4 d[8];
    auto gather = [&](int k0, int id_raw, float v_raw) {
        const int l31 = lane_here() & 31;
        const bool in = k0 + l31 < n;
        ibuf[l31] = in ? id_raw : -1;
        rbuf[l31] = in ? v_raw : 0.f;
        wave_sync();
        if (active) {
            // the task's eight indices in two 16-byte reads, so the loads
            // wait for LDS once rather than once each
            const int gi = lane_here() / nseg;  // = g, recomputed rather than kept
            const int4 i0 = *reinterpret_cast<const int4 *>(ibuf + 8 * gi);
            const int4 i1 = *reinterpret_cast<const int4 *>(ibuf + 8 * gi + 4);
            const int ids[8] = {i0.x, i0.y, i0.z, i0.w, i1.x, i1.y, i1.z, i1.w};
            if (a.ys_bytes > 0) {
                // Buffer loads with a 32-bit byte offset: a slot past the
                // row's end (id -1) has an offset of 2^32 - 2 FP + 2 soff,
                // beyond the buffer, and reads zeros.
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    d[e] = __builtin_bit_cast(
                        u32x4, __builtin_amdgcn_raw_buffer_load_b128(ysr, ids[e] * (2 * G::FP) + soff2, 0, 0));
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const uint16_t *src = (ids[e] >= 0 ? a.Ys + (int64_t)ids[e] * G::FP : a.zero16) + soff;
                    d[e] = *reinterpret_cast<const u32x4 *>(src);
                }
            }
        }
    };
    int nid;
    float nval;
    load_ids(0, nid, nval);
    gather(0, nid, nval);
    load_ids(CHUNK, nid, nval);
    for (int k0 = 0; k0 < n; k0 += CHUNK) {
        if (active) {
#pragma unroll
            for (int e = 0; e < 8; ++e) *reinterpret_cast<u32x4 *>(wbase + e * RS) = d[e];
        }
        if (const int lh = lane_here(); lh < CHUNK) Yr[lh * RS + f] = to16<E>(rbuf[lh]);
        wave_sync();
        u32x4 F[T];
#pragma unroll
        for (int I = 0; I < T; ++I) {
            const v4i16 t0 = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4i16 *)(rbase + 16 * I));
            const v4i16 t1 = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4i16 *)(rbase + 4 * RS + 16 * I));
            const u32x2 w0 = __builtin_bit_cast(u32x2, t0), w1 = __builtin_bit_cast(u32x2, t1);
            F[I] = u32x4{w0[0], w0[1], w1[0], w1[1]};
        }
        if (k0 + CHUNK < n) {
            // The operand reads above are issued before the next chunk's
            // gathers write the stage, and LDS keeps a wave's operations in
            // order; ibuf and rbuf are read before they are rewritten.
            gather(k0 + CHUNK, nid, nval);
            load_ids(k0 + 2 * CHUNK, nid, nval);
        }
#pragma unroll
        for (int I = 0; I < T; ++I)
#pragma unroll
            for (int J = 0; J <= I; ++J) {
                if constexpr (E == PREC_BF16)
                    acc[tile_id(I, J)] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(bf16x8, F[J]), __builtin_bit_cast(bf16x8, F[I]), acc[tile_id(I, J)],
                        0, 0, 0);
                else
                    acc[tile_id(I, J)] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                        __builtin_bit_cast(f16x8, F[J]), __builtin_bit_cast(f16x8, F[I]), acc[tile_id(I, J)],
                        0, 0, 0);
            }
        wave_sync();
    }
}

// 16-bit Gram of one row with the gathers going straight to LDS (LDS-DMA:
// buffer_load_dwordx4 ... lds, no registers), two chunks in flight per wave.
// Lane (g, s) of the 4 * 2T active lanes loads segment s (factors 8s ..
// 8s+7) of ratings 4e + g, e = 0..7; load e of the wave writes its active
// lanes' 16 bytes at 16 * lane from buffer row 4e, i.e. rating 4e + g at row
// 4e + g, column 8s: the rating-major stage of gram_row16_piped, filled in
// rating order, so the transposing operand reads and the MFMAs are the
// same as there and the sums the same bit for bit. Slots past the row's end
// are zeroed in LDS after their chunk has landed.
// The order of vector-memory operations is fixed and counted: per chunk j,
// I(j) = the index and rating loads (two buffer loads) and D(j) = its eight
// LDS-DMA loads, issued as I0 I1 D0 I2 D1 I3, then per chunk k consumed:
// D(k+2) and I(k+4), always (past the row's end they touch no memory), so
// that when chunk k is consumed the operations after D(k) are I(k+2),
// D(k+1) and I(k+3): s_waitcnt vmcnt(12) retires D(k) by count. The
// compiler adds its own wait before the operand reads, for every LDS-DMA
// load in flight (it cannot tell the two buffers apart), and the loop
// relies on that wider wait too: with the count alone, the sums were wrong
// at T = 3, where a step is shortest.
// Before a buffer is refilled, lgkmcnt(0) makes sure its operand reads are
// done. Index and rating loads past the row's end read its last rating.
// The loop of gram_row16_dma, with the two stage buffers and the two index /
// rating buffer pairs as distinct pointers.
template <int T, int E>
__device__ __forceinline__ void gram_dma_loop(const UpdateArgs &a, int64_t rs, int n, f32x4 (&acc)[Geo<T, E>::NT],
                                              uint16_t *__restrict__ buf0, uint16_t *__restrict__ buf1,
                                              int *__restrict__ ib0, int *__restrict__ ib1,
                                              float *__restrict__ rb0, float *__restrict__ rb1) {
    using G = Geo<T, E>;
    constexpr int RS = G::FP;             // stage row stride, halves
    constexpr int NS = 2 * T;             // 16-byte segments per row
    typedef short v4i16 __attribute__((ext_vector_type(4)));
    typedef __attribute__((address_space(3))) v4i16 lds_v4i16;
    typedef __attribute__((address_space(3))) void lds_void;
    const int l = lane();
    const int f = a.f;
    const bool active = l < 4 * NS;
    const int g = l / NS, s = l - g * NS;
    const int soff2 = 16 * s;
    const int roff = (8 * (l >> 4) + ((l & 15) >> 2)) * RS + 4 * (l & 3);
    const int nch = (n + CHUNK - 1) / CHUNK;
    const __amdgpu_buffer_rsrc_t ysr =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t *>(a.Ys), (short)0, a.ys_bytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t isr =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<int *>(a.idx + rs), (short)0, n * 4, 0x00020000);
    const __amdgpu_buffer_rsrc_t vsr =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(a.val + rs), (short)0, n * 4, 0x00020000);
    // I(j): one buffer load each for the index and the rating of slot l % 32
    auto load_ids = [&](int j, int &id_raw, float &v_raw) {
        const int off = 4 * min(CHUNK * j + (lane_here() & 31), n - 1);
        id_raw = __builtin_amdgcn_raw_buffer_load_b32(isr, off, 0, 0);
        v_raw = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(vsr, off, 0, 0));
    };
    // publish chunk j (parity P): ib[8g + e] = index of rating 4e + g (-1
    // past the row's end), rb[m] = rating m
    auto publish = [&](auto parity, int j, int id_raw, float v_raw) {
        constexpr int P = decltype(parity)::value;
        int *ib = P ? ib1 : ib0;
        float *rb = P ? rb1 : rb0;
        const int m = lane_here() & 31;
        const bool in = CHUNK * j + m < n;
        ib[8 * (m & 3) + (m >> 2)] = in ? id_raw : -1;
        rb[m] = in ? v_raw : 0.f;
    };
    // D(j) (parity P): the eight LDS-DMA loads of chunk j into buffer P
    auto dma = [&](auto parity) {
        constexpr int P = decltype(parity)::value;
        const int *ib = P ? ib1 : ib0;
        uint16_t *dst = P ? buf1 : buf0;
        wave_sync();
        if (active) {
            const int gi = lane_here() / NS;  // = g, recomputed rather than kept
            const int4 i0 = *reinterpret_cast<const int4 *>(ib + 8 * gi);
            const int4 i1 = *reinterpret_cast<const int4 *>(ib + 8 * gi + 4);
            const int ids[8] = {i0.x, i0.y, i0.z, i0.w, i1.x, i1.y, i1.z, i1.w};
#pragma unroll
            for (int e = 0; e < 8; ++e)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(ysr, (lds_void *)(dst + 4 * e * RS), 16,
                                                         ids[e] * (2 * G::FP) + soff2, 0, 0, 0);
        }
    };
    using P0 = std::integral_constant<int, 0>;
    using P1 = std::integral_constant<int, 1>;
    int id0, id1;
    float v0, v1;
    load_ids(0, id0, v0);
    load_ids(1, id1, v1);
    publish(P0{}, 0, id0, v0);
    dma(P0{});
    load_ids(2, id0, v0);
    publish(P1{}, 1, id1, v1);
    dma(P1{});
    load_ids(3, id1, v1);
    // Consumes chunk k (parity P = k % 2, static, so the index and rating
    // registers it publishes are known at compile time) and issues D(k+2)
    // and I(k+4), unconditionally: a chunk past the row's end has every
    // slot out of range (index -1), so its loads touch no memory.
    auto step = [&](auto parity, int k) {
        constexpr int P = decltype(parity)::value;
        // D(k) has landed: after it come I(k+2), D(k+1), I(k+3) (12 loads)
        asm volatile("s_waitcnt vmcnt(12)" ::: "memory");
        uint16_t *buf = P ? buf1 : buf0;
        const float *rb = P ? rb1 : rb0;
        const int nv = min(CHUNK, n - CHUNK * k);
        if (nv < CHUNK) {
            // rows past the row's end: zero (16-byte pieces; piece p is row
            // nv + p / NS, segment p % NS)
            for (int p = l; p < (CHUNK - nv) * NS; p += WAVE)
                *reinterpret_cast<u32x4 *>(buf + (nv + p / NS) * RS + 8 * (p % NS)) = u32x4{0u, 0u, 0u, 0u};
        }
        if (const int lh = lane_here(); lh < CHUNK) buf[lh * RS + f] = to16<E>(rb[lh]);
        wave_sync();
        u32x4 F[T];
#pragma unroll
        for (int I = 0; I < T; ++I) {
            const v4i16 t0 = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4i16 *)(buf + roff + 16 * I));
            const v4i16 t1 = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4i16 *)(buf + roff + 4 * RS + 16 * I));
            const u32x2 w0 = __builtin_bit_cast(u32x2, t0), w1 = __builtin_bit_cast(u32x2, t1);
            F[I] = u32x4{w0[0], w0[1], w1[0], w1[1]};
        }
        // chunk k + 2 into this buffer once its operand reads are done
        if constexpr (P == 0) publish(P0{}, k + 2, id0, v0);
        else publish(P1{}, k + 2, id1, v1);
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        dma(parity);
        if constexpr (P == 0) load_ids(k + 4, id0, v0);
        else load_ids(k + 4, id1, v1);
#pragma unroll
        for (int I = 0; I < T; ++I)
#pragma unroll
            for (int J = 0; J <= I; ++J) {
                if constexpr (E == PREC_BF16)
                    acc[tile_id(I, J)] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(bf16x8, F[J]), __builtin_bit_cast(bf16x8, F[I]), acc[tile_id(I, J)],
                        0, 0, 0);
                else
                    acc[tile_id(I, J)] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                        __builtin_bit_cast(f16x8, F[J]), __builtin_bit_cast(f16x8, F[I]), acc[tile_id(I, J)],
                        0, 0, 0);
            }
    };
    int k = 0;
    for (; k + 1 < nch; k += 2) {
        step(P0{}, k);
        step(P1{}, k + 1);
    }
    if (k < nch) step(P0{}, k);
    // every load issued here has landed before the stage is used again
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

template <int T, int E>
__device__ __forceinline__ void gram_row16_dma(const UpdateArgs &a, int64_t rs, int n,
                                               f32x4 (&acc)[Geo<T, E>::NT], char *lds) {
    using G = Geo<T, E>;
    constexpr int BUF = CHUNK * G::FP;  // halves per buffer
    if (n <= 0) return;
    uint16_t *stage = reinterpret_cast<uint16_t *>(lds + G::MISC_FLOATS * 4);
    gram_dma_loop<T, E>(a, rs, n, acc, stage, stage + BUF, reinterpret_cast<int *>(lds),
                        reinterpret_cast<int *>(stage + 2 * BUF), reinterpret_cast<float *>(lds) + 32,
                        reinterpret_cast<float *>(stage + 2 * BUF) + 32);
}

template <int T, int E>
__device__ __forceinline__ void gram_row(const UpdateArgs &a, int64_t rs, int n,
                                         f32x4 (&acc)[Geo<T, E>::NT], char *lds) {
    using G = Geo<T, E>;
    const int l = lane();
    char *stage = lds + G::MISC_FLOATS * 4;
#pragma unroll
    for (int t = 0; t < G::NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    if constexpr (G::DMA_GRAM) {
        if (a.ys_bytes > 0 && a.gram_dma) {
            gram_row16_dma<T, E>(a, rs, n, acc, lds);
            return;
        }
    }
    if constexpr (E != PREC_FP32 && T <= 9) {
        // f <= 128: at most 16 segments, one task per lane (every f of
        // T <= 8; f = 128 of T = 9, whose rating column is the 129th)
        if (T <= 8 || a.f <= 128) {
            gram_row16_piped<T, E>(a, rs, n, acc, lds);
            return;
        }
    }
    stage_clear<T, E>(a.f, stage);
    for (int k0 = 0; k0 < n; k0 += CHUNK) {
        stage_chunk<T, E>(a, rs, k0, min(CHUNK, n - k0), lds);
        if constexpr (E == PREC_FP32) {
            const float *Ys = reinterpret_cast<const float *>(stage);
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
                            __builtin_amdgcn_mfma_f32_16x16x4f32(F[J], F[I], acc[tile_id(I, J)], 0, 0, 0);
            }
        } else {
            const uint16_t *Yt = reinterpret_cast<const uint16_t *>(stage);
            u32x4 F[T];
#pragma unroll
            for (int I = 0; I < T; ++I)
                F[I] = *reinterpret_cast<const u32x4 *>(Yt + G::yt_row(16 * I + (l & 15)) * G::YT_H + 8 * (l >> 4));
#pragma unroll
            for (int I = 0; I < T; ++I)
#pragma unroll
                for (int J = 0; J <= I; ++J) {
                    if constexpr (E == PREC_BF16)
                        acc[tile_id(I, J)] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            __builtin_bit_cast(bf16x8, F[J]), __builtin_bit_cast(bf16x8, F[I]), acc[tile_id(I, J)],
                            0, 0, 0);
                    else
                        acc[tile_id(I, J)] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                            __builtin_bit_cast(f16x8, F[J]), __builtin_bit_cast(f16x8, F[I]), acc[tile_id(I, J)],
                            0, 0, 0);
                }
        }
        wave_sync();
    }
}

// Diagonal: + lambda * n on the f real pivots; padding pivots (index > f)
// become 1 so the padded system is the identity there. Index f (the rating
// column's own square sum) is left alone: it is never a pivot.
template <int T, int NT>
__device__ __forceinline__ void add_diag(f32x4 (&acc)[NT], int f, float reg) {
    int l = lane();
    // Opaque per row: the lane masks stay per-row vector compares instead of
    // being hoisted out of the row loop into scalar registers (which spill).
    asm volatile("" : "+v"(l));
    const int q = l >> 4, c = l & 15, rr = c & 3;
    // Lane (q, c) holds diagonal entry c of every diagonal tile when
    // c / 4 == q, in register c % 4.
    const bool mine = (c >> 2) == q;
#pragma unroll
    for (int I = 0; I < T; ++I) {
        const int d = 16 * I + c;
        f32x4 &t = acc[tile_id(I, I)];
        const float v = pick4(t, rr);
        const float nv = d < f ? v + reg : d > f ? 1.f : v;
#pragma unroll
        for (int r = 0; r < 4; ++r) t[r] = (mine && rr == r) ? nv : t[r];
    }
}

// ------------------------------------------------------- lane exchange
// A 64-lane wave is 4 rows (q = lane / 16) of 16 lanes (c = lane % 16).

// DPP moves for patterns whose source lanes are all valid (quad_perm,
// row_mirror, row_ror): bound_ctrl set, so no "old" value has to be
// materialised in the destination first.
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 16 lanes of a row. Quad sums, then mirror and rotate-by-8
// pairings: every lane adds the same operands in the same order, so all 16
// lanes hold identical bits.
__device__ __forceinline__ float row16_sum(float v) {
    v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_mov<0x140>(v);  // row_mirror
    v += dpp_mov<0x128>(v);  // row_ror:8
    return v;
}
__device__ __forceinline__ double row16_sum_d(double v) {
    v += dpp_mov_d<0xB1>(v);
    v += dpp_mov_d<0x4E>(v);
    v += dpp_mov_d<0x140>(v);
    v += dpp_mov_d<0x128>(v);
    return v;
}

// Row G's value at this lane's column, in every row. permlane16_swap(v, v)
// leaves the even row of each row pair in the first result and the odd row
// in the second; permlane32_swap(t, t) does the same for the two halves.
template <int G>
__device__ __forceinline__ float rows_bcast(float v) {
    const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    const unsigned int t = (G & 1) ? p[1] : p[0];
    const auto s = __builtin_amdgcn_permlane32_swap(t, t, false, false);
    return __uint_as_float((G & 2) ? s[1] : s[0]);
}

// Sum over the 4 rows: permlane16 swap pairs rows (0,1), (2,3); permlane32
// swap pairs the two halves. Each pair sum is formed from the same two
// operands in every lane.
__device__ __forceinline__ float q4_sum(float v) {
    const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    v = __uint_as_float(p[0]) + __uint_as_float(p[1]);
    const auto s = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(s[0]) + __uint_as_float(s[1]);
}
__device__ __forceinline__ double q4_sum_d(double v) {
    const long long b = __double_as_longlong(v);
    const auto plo = __builtin_amdgcn_permlane16_swap((unsigned int)b, (unsigned int)b, false, false);
    const auto phi = __builtin_amdgcn_permlane16_swap((unsigned int)(b >> 32), (unsigned int)(b >> 32), false, false);
    v = __longlong_as_double(((long long)phi[0] << 32) | plo[0]) + __longlong_as_double(((long long)phi[1] << 32) | plo[1]);
    const long long b2 = __double_as_longlong(v);
    const auto slo = __builtin_amdgcn_permlane32_swap((unsigned int)b2, (unsigned int)b2, false, false);
    const auto shi = __builtin_amdgcn_permlane32_swap((unsigned int)(b2 >> 32), (unsigned int)(b2 >> 32), false, false);
    return __longlong_as_double(((long long)shi[0] << 32) | slo[0]) + __longlong_as_double(((long long)shi[1] << 32) | slo[1]);
}

// ------------------------------------------------------- vector forms
// A vector v of the padded dimension 16T lives in one of two register forms:
//   column form  vc[J]      = v[16J + c]          (replicated over q)
//   row form     vr[J][r]   = v[16J + 4q + r]     (replicated over c)

__device__ __forceinline__ f32x4 col_to_row(float vcJ) {
    const int l = lane(), q = l >> 4;
    f32x4 o;
#pragma unroll
    for (int r = 0; r < 4; ++r) o[r] = __shfl(vcJ, (l & 48) | (4 * q + r), WAVE);
    return o;
}
__device__ __forceinline__ double col_to_row_d(double vcJ, int r) {
    const int l = lane(), q = l >> 4;
    return __shfl(vcJ, (l & 48) | (4 * q + r), WAVE);
}

// Lane (q, c) takes v[c] from row c/4, where every lane holds the row's four
// entries; the source lane picks the entry the destination needs.
__device__ __forceinline__ float row_to_col(const f32x4 &vrJ) {
    const int c = lane() & 15;
    return __shfl(pick4(vrJ, c & 3), ((c >> 2) << 4) | c, WAVE);
}

// ------------------------------------------------------- tile layout
// The augmented Gram is held as its lower 16x16 tiles, a[tile_id(I,J)],
// I >= J, in MFMA accumulator registers (lane (q,c), register r):
//   diagonal tile   (I,I): A[16I + 4q + r][16I + c]  (full, symmetric)
//   off-diagonal    (I,J): A[16I + c][16J + 4q + r]  ("row by lane")
// The Gram produces off-diagonal tiles in this second layout directly (as
// the transposed upper tile). With it, every product the blocked Cholesky
// needs takes its MFMA operands straight from registers.
//
// After chol_regs, off-diagonal tiles hold L (same layout) and each
// diagonal tile holds W = L_KK^-1 in the row-by-lane layout:
// a[tile_id(K,K)][r] at lane (q,c) = W[c][4q + r]. Row f of the last block
// row (the augmented row) holds z = L^-1 b for blocks K < T-1.

// Row f of the augmented matrix before the factorisation (b, masked to
// the first f entries), column form.
template <int T, int NT>
__device__ __forceinline__ void extract_b(const f32x4 (&a)[NT], int f, float (&bc)[T]) {
    const int l = lane(), c = l & 15;
#pragma unroll
    for (int K = 0; K < T; ++K) {
        // lane (q, f & 15) holds A[f][16K + 4q + r] (off-diagonal layout, or
        // column f of the symmetric diagonal tile)
        f32x4 br;
#pragma unroll
        for (int r = 0; r < 4; ++r) br[r] = __shfl(a[tile_id(T - 1, K)][r], (l & 48) | (f & 15), WAVE);
        const float v = row_to_col(br);
        bc[K] = (16 * K + c < f) ? v : 0.f;
    }
}

// ------------------------------------------------------------- Cholesky
// Step J of the diagonal-block factorisation D = L~ P L~^T (L~ unit lower,
// P the pivots, L = L~ P^1/2). D is in the standard layout, kept symmetric;
// the step takes pivot J, applies the rank-1 update with coefficient
// -D[J][c] / P[J] to the columns past J, and the matching forward-elimination
// step of W~ = L~^-1 (W~[c] -= L~[c][J] W~[J], c > J), W~ held row by lane.
// Both updates share the coefficient, and W~'s rows need no scaling until
// the end (W = P^-1/2 W~: row c times 1/sqrt(P[c])).
// The rank-1 update is applied to every row of the columns past J: rows of
// finished pivots i <= J hold the upper-triangle entries D[i][c], which the
// same symmetric update drives to (rounding-level) zero and which no later
// step reads for a result. A pivot <= 0 makes its 1/sqrt non-finite, which
// reaches x through W; the caller's finiteness check fails the row.
template <int J, bool FULL>
__device__ __forceinline__ void diag_steps(f32x4 &d, f32x4 &w, int jend) {
    if constexpr (J < 16) {
        if (FULL || J < jend) {
            int c = lane() & 15;
            // Opaque per step: the lane masks below are recomputed here
            // instead of being hoisted for all 16 steps at once (which runs
            // out of scalar registers).
            asm volatile("" : "+v"(c));
            constexpr int JG = J >> 2, JR = J & 3;
            const float piv = bcast_lane(d[JR], (JG << 4) | J);
            const float rs = __builtin_amdgcn_rsqf(piv);
            // The broadcast runs in every lane (not under the c > J mask), so
            // the permlane swaps see a full wave and need no branch.
            const float rb = rows_bcast<JG>(d[JR]);
            const float coef = (c > J) ? -rb * (rs * rs) : 0.f;  // -D[J][c] / P[J]
            // d[r] += bcastJ(d[r]) * coef   (-= D[4q+r][J] D[J][c] / P[J])
            // w[r] += bcastJ(w[r]) * coef   (-= L~[c][J] W~[J][4q+r])
            // bcastJ(v): v of lane J of this lane's row (DPP row_newbcast:J)
            // with the broadcast folded into the FMA as its DPP source operand:
            // one instruction per element instead of a DPP move plus an FMA.
            // s_nop 1 gives the two wait states a DPP read needs after a VALU
            // write of its source; inside the block no instruction reads the
            // register its predecessor wrote.
            asm volatile(
                "s_nop 1\n\t"
                "v_fmac_f32_dpp %0, %0, %8 row_newbcast:%9 row_mask:0xf bank_mask:0xf\n\t"
                "v_fmac_f32_dpp %1, %1, %8 row_newbcast:%9 row_mask:0xf bank_mask:0xf\n\t"
                "v_fmac_f32_dpp %2, %2, %8 row_newbcast:%9 row_mask:0xf bank_mask:0xf\n\t"
                "v_fmac_f32_dpp %3, %3, %8 row_newbcast:%9 row_mask:0xf bank_mask:0xf\n\t"
                "v_fmac_f32_dpp %4, %4, %8 row_newbcast:%9 row_mask:0xf bank_mask:0xf\n\t"
                "v_fmac_f32_dpp %5, %5, %8 row_newbcast:%9 row_mask:0xf bank_mask:0xf\n\t"
                "v_fmac_f32_dpp %6, %6, %8 row_newbcast:%9 row_mask:0xf bank_mask:0xf\n\t"
                "v_fmac_f32_dpp %7, %7, %8 row_newbcast:%9 row_mask:0xf bank_mask:0xf"
                : "+v"(d[0]), "+v"(d[1]), "+v"(d[2]), "+v"(d[3]), "+v"(w[0]), "+v"(w[1]), "+v"(w[2]), "+v"(w[3])
                : "v"(coef), "i"(J));
        }
        diag_steps<J + 1, FULL>(d, w, jend);
    }
}

// Factor the diagonal block with jend pivots (FULL: jend == 16); on return
// the tile holds W = L^-1 row by lane, rows >= jend zero.
template <bool FULL>
__device__ __forceinline__ void diag_factor(f32x4 &d, int jend) {
    const int l = lane(), q = l >> 4, c = l & 15;
    f32x4 w;
#pragma unroll
    for (int r = 0; r < 4; ++r) w[r] = (c == 4 * q + r) ? 1.f : 0.f;
    diag_steps<0, FULL>(d, w, jend);
    // Row scaling 1/sqrt(P[c]). Steps past c leave column c alone (their
    // coefficient is zero there), so D[c][c] is still the value step c
    // pivoted on; it sits in row c/4, register c%4, and the sum over the
    // four rows adds three zeros to it.
    const float dc = pick4(d, c & 3);
    const float myrs = __builtin_amdgcn_rsqf(q4_sum((q == (c >> 2)) ? dc : 0.f));
#pragma unroll
    for (int r = 0; r < 4; ++r) d[r] = (FULL || c < jend) ? w[r] * myrs : 0.f;
}

// A tile's four registers as FP16 hi and lo parts, x = hi + lo + O(2^-22 x):
// the operand of a 16x16x16 FP16 MFMA in the register layout of the FP32
// 16x16x4 form (lane (q, c) holds k = 4q .. 4q+3 of row c). hi = fp16(x)
// and lo = fp16(x - hi), both rounded to nearest; x - hi is exact in FP32.
// Two instructions per value: packed conversions (two values each) and
// v_fma_mix_f32, which reads hi as FP16 and forms x - hi in one step.
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
template <int HALF>
__device__ __forceinline__ float minus_half(f16x2 h, float x) {  // x - h[HALF], exact
    float r;
    if constexpr (HALF == 0)
        asm("v_fma_mix_f32 %0, %1, -1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h), "v"(x));
    else
        asm("v_fma_mix_f32 %0, %1, -1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h), "v"(x));
    return r;
}
__device__ __forceinline__ void split16(const f32x4 &x, f16x4 &hi, f16x4 &lo) {
    const f16x2 h01 = __builtin_convertvector(f32x2{x[0], x[1]}, f16x2);
    const f16x2 h23 = __builtin_convertvector(f32x2{x[2], x[3]}, f16x2);
    const f16x2 l01 = __builtin_convertvector(f32x2{minus_half<0>(h01, x[0]), minus_half<1>(h01, x[1])}, f16x2);
    const f16x2 l23 = __builtin_convertvector(f32x2{minus_half<0>(h23, x[2]), minus_half<1>(h23, x[3])}, f16x2);
    hi = __builtin_shufflevector(h01, h23, 0, 1, 2, 3);
    lo = __builtin_shufflevector(l01, l23, 0, 1, 2, 3);
}

// Blocked right-looking Cholesky, panel K:
//   diagonal block -> W_KK = L_KK^-1 (diag_factor)
//   panel          L_IK^T = W_KK A_IK^T       (4 MFMAs per tile)
//   trailing       A_IJ  -= L_IK L_JK^T        (4 FP32 MFMAs per tile, 128
//                  cycles; with SPLIT two FP16 16x16x32 MFMAs on the hi/lo
//                  parts of the panel tiles, 32 cycles, each FP32 product
//                  kept to ~2^-22 relative; the entries of L are bounded by
//                  sqrt(A_ii), inside FP16's range for the ratings
//                  als_set_train accepts)
// The last block holds row f (the augmented row, never a pivot): yl gets
// its reduced entries (row form) just before that block is factored.
template <int T, int K, int NT, bool SPLIT>
__device__ __forceinline__ void chol_panels(f32x4 (&a)[NT], int f, f32x4 &yl) {
    if constexpr (K < T) {
        const int jend = min(16, f - 16 * K);
        if constexpr (K == T - 1) {
            const int l = lane();
#pragma unroll
            for (int r = 0; r < 4; ++r) yl[r] = __shfl(a[tile_id(K, K)][r], (l & 48) | (f & 15), WAVE);
        }
        diag_factor<(K + 1 < T)>(a[tile_id(K, K)], jend);
        if constexpr (K + 1 < T) {
#pragma unroll
            for (int I = K + 1; I < T; ++I) {
                const f32x4 o = a[tile_id(I, K)];
                f32x4 nw = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    nw = __builtin_amdgcn_mfma_f32_16x16x4f32(a[tile_id(K, K)][r], o[r], nw, 0, 0, 0);
                a[tile_id(I, K)] = nw;
            }
            if constexpr (SPLIT) {
                // Each panel tile as [hi | lo], the 8 halves a lane feeds a
                // 16x16x32 MFMA (the contraction pairs position j of lane
                // (q, c) in one operand with position j of lane (q, c') in the
                // other). Against [hi_J | hi_J] it gives hi_J hi_I + hi_J lo_I,
                // against [lo_J | lo_J] lo_J hi_I + lo_J lo_I: the whole
                // product in two MFMAs of 16 cycles (three 16x16x16 ones
                // took 48).
                f16x8 hl[T];
#pragma unroll
                for (int I = K + 1; I < T; ++I) {
                    f16x4 h, l;
                    split16(a[tile_id(I, K)], h, l);
                    hl[I] = __builtin_shufflevector(h, l, 0, 1, 2, 3, 4, 5, 6, 7);
                }
#pragma unroll
                for (int J = K + 1; J < T; ++J) {
                    const f16x8 nhh = -__builtin_shufflevector(hl[J], hl[J], 0, 1, 2, 3, 0, 1, 2, 3);
                    const f16x8 nll = -__builtin_shufflevector(hl[J], hl[J], 4, 5, 6, 7, 4, 5, 6, 7);
#pragma unroll
                    for (int I = J; I < T; ++I) {
                        f32x4 t = a[tile_id(I, J)];
                        t = __builtin_amdgcn_mfma_f32_16x16x32_f16(nhh, hl[I], t, 0, 0, 0);
                        t = __builtin_amdgcn_mfma_f32_16x16x32_f16(nll, hl[I], t, 0, 0, 0);
                        a[tile_id(I, J)] = t;
                    }
                }
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int J = K + 1; J < T; ++J) {
                        const float nf = -a[tile_id(J, K)][r];
#pragma unroll
                        for (int I = J; I < T; ++I)
                            a[tile_id(I, J)] = __builtin_amdgcn_mfma_f32_16x16x4f32(nf, a[tile_id(I, K)][r],
                                                                                   a[tile_id(I, J)], 0, 0, 0);
                    }
            }
        }
        chol_panels<T, K + 1, NT, SPLIT>(a, f, yl);
    }
}

// Factorises the augmented matrix in place and returns z = L^-1 b in row
// form (zero past f). A pivot <= 0 leaves non-finite values that reach x
// (the caller checks x); the return value is kept for that interface.
// SPLIT: the trailing updates on FP16 hi/lo parts (chol_panels).
template <int T, int NT, bool SPLIT = false>
__device__ __forceinline__ bool chol_regs(f32x4 (&a)[NT], int f, f32x4 (&zr)[T]) {
    f32x4 yl;
    chol_panels<T, 0, NT, SPLIT>(a, f, yl);
    const int l = lane(), q = l >> 4, c = l & 15;
#pragma unroll
    for (int K = 0; K + 1 < T; ++K)
#pragma unroll
        for (int r = 0; r < 4; ++r) zr[K][r] = __shfl(a[tile_id(T - 1, K)][r], (l & 48) | (f & 15), WAVE);
    // last block: z = W y (W rows past the pivots are zero)
    const int jend = f - 16 * (T - 1);
    float u = 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) u += a[tile_id(T - 1, T - 1)][r] * ((4 * q + r < jend) ? yl[r] : 0.f);
    const float zc = q4_sum(u);
    zr[T - 1] = col_to_row(c < jend ? zc : 0.f);
    return true;
}

// -------------------------------------------------- triangular solves
// L^T x = z. z in row form; x out in column form (zero past f).
//   x_K = W_KK^T (z_K - sum_{I>K} L_IK^T x_I)
template <int T, int NT>
__device__ __forceinline__ void back_solve(const f32x4 (&a)[NT], int f, const f32x4 (&zr)[T], float (&xc)[T]) {
    const int c = lane() & 15;
#pragma unroll
    for (int K = T - 1; K >= 0; --K) {
        f32x4 p = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int I = K + 1; I < T; ++I) p += a[tile_id(I, K)] * xc[I];
        f32x4 t;
#pragma unroll
        for (int r = 0; r < 4; ++r) t[r] = (K + 1 < T) ? zr[K][r] - row16_sum(p[r]) : zr[K][r];
        const float tc = row_to_col(t);
        f32x4 xr;
#pragma unroll
        for (int r = 0; r < 4; ++r) xr[r] = row16_sum(a[tile_id(K, K)][r] * tc);
        const float v = row_to_col(xr);
        xc[K] = (16 * K + c < f) ? v : 0.f;
    }
}

// L w = v. v in column form (entries past f are ignored); w out in row form
// (zero past f).
//   w_K = W_KK (v_K - sum_{J<K} L_KJ w_J)
template <int T, int NT>
__device__ __forceinline__ void fwd_solve(const f32x4 (&a)[NT], int f, const float (&vc)[T], f32x4 (&wr)[T]) {
    const int q = lane() >> 4;
#pragma unroll
    for (int K = 0; K < T; ++K) {
        float s = 0.f;
#pragma unroll
        for (int J = 0; J < K; ++J)
#pragma unroll
            for (int r = 0; r < 4; ++r) s += a[tile_id(K, J)][r] * wr[J][r];
        const float tc = (K > 0) ? vc[K] - q4_sum(s) : vc[K];
        const f32x4 tr = col_to_row(tc);
        float u = 0.f;
#pragma unroll
        for (int r = 0; r < 4; ++r) u += a[tile_id(K, K)][r] * tr[r];
        const f32x4 w = col_to_row(q4_sum(u));
#pragma unroll
        for (int r = 0; r < 4; ++r) wr[K][r] = (16 * K + 4 * q + r < f) ? w[r] : 0.f;
    }
}

// ------------------------------------------------------ refinement
// Iterative refinement, x += A^-1 (b - A x): the residual in FP64, the
// correction solved on the FP32 factor in the registers.

template <int G>
__device__ __forceinline__ double rows_bcast_d(double v) {
    const long long b = __double_as_longlong(v);
    const unsigned int lo = __float_as_uint(rows_bcast<G>(__uint_as_float((unsigned int)b)));
    const unsigned int hi = __float_as_uint(rows_bcast<G>(__uint_as_float((unsigned int)(b >> 32))));
    return __longlong_as_double(((long long)hi << 32) | lo);
}

// Sum of the two 32-lane halves, in every lane.
__device__ __forceinline__ double halves_sum_d(double v) {
    const long long b = __double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned int)b, (unsigned int)b, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned int)(b >> 32), (unsigned int)(b >> 32), false, false);
    return __longlong_as_double(((long long)hi[0] << 32) | lo[0]) + __longlong_as_double(((long long)hi[1] << 32) | lo[1]);
}

// Sums of four values over the 16 lanes of a row, scattered: lane c gets
// the total of v[c & 3]. Two exchange levels inside the quad (partners c^2,
// then c^1) halve the values carried, then the four quads are added with
// rotations by 4 and 8, which keep c & 3. Lanes c, c+4, c+8, c+12 hold the
// same total, possibly rounded differently; callers read one fixed lane.
__device__ __forceinline__ double rows16_scatter_d(const double (&v)[4]) {
    const int c = lane() & 15;
    const bool b1 = (c & 2) != 0, b0 = (c & 1) != 0;
    const double k0 = b1 ? v[2] : v[0], k1 = b1 ? v[3] : v[1];
    const double s0 = b1 ? v[0] : v[2], s1 = b1 ? v[1] : v[3];
    const double u0 = k0 + dpp_mov_d<0x4E>(s0);  // quad_perm [2,3,0,1]
    const double u1 = k1 + dpp_mov_d<0x4E>(s1);
    double t = (b0 ? u1 : u0) + dpp_mov_d<0xB1>(b0 ? u0 : u1);  // quad_perm [1,0,3,2]
    t += dpp_mov_d<0x124>(t);  // row_ror:4
    t += dpp_mov_d<0x128>(t);  // row_ror:8
    return t;
}

// Warm start: replaces the augmented row f (b) by r0 = b - A x0, formed in
// FP64 from the tiles still in the registers (before the factorisation), so
// that the factorisation yields L^-1 r0 in that row. Block K of y = A x0:
//   tiles (K, J <= K) give row 16K + c from lane (q, c) (summed over q),
//   tiles (I > K, K) give row 16K + 4q + r from register r (summed over c).
// b is copied to bs (16T floats of LDS) before any block is overwritten;
// row f enters the other blocks' sums only through x0[f] = 0 or through the
// discarded entry f. x0 in column form (zero past f); xs (16T floats of LDS)
// holds it in memory order, where the row form of a block is one 16-byte
// read. The last diagonal tile keeps its row and column f symmetric.
template <int T, int NT>
__device__ __forceinline__ void warm_row_f(f32x4 (&a)[NT], int f, const float (&xc)[T], float *xs, float *bs) {
    const int l = lane(), q = l >> 4, c = l & 15, fr = f & 15;
    if (l < 16) {
#pragma unroll
        for (int J = 0; J < T; ++J) xs[16 * J + l] = xc[J];
    }
    if (c == fr) {
        // lane (q, f % 16) holds b[16K + 4q + r] in register r of tile (T-1, K)
#pragma unroll
        for (int K = 0; K < T; ++K) *reinterpret_cast<f32x4 *>(bs + 16 * K + 4 * q) = a[tile_id(T - 1, K)];
    }
    wave_sync();
#pragma unroll
    for (int K = 0; K < T; ++K) {
        double cp = 0.0;
#pragma unroll
        for (int J = 0; J <= K; ++J) {
            const f32x4 xJ = *reinterpret_cast<const f32x4 *>(xs + 16 * J + 4 * q);
#pragma unroll
            for (int r = 0; r < 4; ++r) cp += (double)a[tile_id(K, J)][r] * (double)xJ[r];
        }
        double rp[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int I = K + 1; I < T; ++I)
#pragma unroll
            for (int r = 0; r < 4; ++r) rp[r] += (double)a[tile_id(I, K)][r] * (double)xc[I];
        // row 16K + c: the column-form part, plus the row-form part taken from
        // lane (c / 4, c), which holds the total for register c % 4 of row c / 4
        const double rs = (K + 1 < T) ? rows16_scatter_d(rp) : 0.0;
        const double y = q4_sum_d(cp) + __shfl(rs, ((c >> 2) << 4) | c, WAVE);
        const float resc = (16 * K + c < f) ? (float)((double)bs[16 * K + c] - y) : 0.f;
        const f32x4 res = col_to_row(resc);  // lane (q, c): r0[16K + 4q + r]
        if (K + 1 < T) {
#pragma unroll
            for (int r = 0; r < 4; ++r) a[tile_id(T - 1, K)][r] = (c == fr) ? res[r] : a[tile_id(T - 1, K)][r];
        } else {
            constexpr int D = tile_id(T - 1, T - 1);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float v = (c == fr) ? res[r] : a[D][r];  // column f
                v = (4 * q + r == fr) ? resc : v;        // row f
                a[D][r] = v;
            }
        }
    }
}

// Residual from this wave's scratch copy Ag of the augmented tiles, taken
// before the factorisation (row f holds b). x column form; r column form
// (zero past f).
template <int T>
__device__ __forceinline__ void residual_copy(const float *Ag, int f, const float (&xc)[T], float (&rc)[T]) {
    const int l = lane(), q = l >> 4;
    f32x4 xr[T];
#pragma unroll
    for (int K = 0; K < T; ++K) xr[K] = col_to_row(xc[K]);
#pragma unroll
    for (int K = 0; K < T; ++K) {
        // Keep this block's scratch loads from being hoisted into earlier
        // blocks (that would hold every tile of A in registers at once).
        __builtin_amdgcn_sched_barrier(0);
        // (A x)_K in column form: sum_{J<=K} A[16K+c][16J+4q+r] x[16J+4q+r]
        // (the diagonal tile is symmetric) ...
        double cp = 0.0;
#pragma unroll
        for (int J = 0; J <= K; ++J) {
            const f32x4 t = *reinterpret_cast<const f32x4 *>(Ag + tile_id(K, J) * 256 + 4 * l);
#pragma unroll
            for (int r = 0; r < 4; ++r) cp += (double)t[r] * (double)xr[J][r];
        }
        // ... plus sum_{I>K} A[16I+c'][16K+4q+r] x[16I+c'] in row form.
        double rp[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int I = K + 1; I < T; ++I) {
            const f32x4 t = *reinterpret_cast<const f32x4 *>(Ag + tile_id(I, K) * 256 + 4 * l);
#pragma unroll
            for (int r = 0; r < 4; ++r) rp[r] += (double)t[r] * (double)xc[I];
        }
        const double cs = q4_sum_d(cp);
        // b_K in row form: lane (q, f & 15) of tile (T-1, K), register r.
        const f32x4 br = *reinterpret_cast<const f32x4 *>(Ag + tile_id(T - 1, K) * 256 + 4 * ((q << 4) | (f & 15)));
        f32x4 res;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const double y = col_to_row_d(cs, r) + row16_sum_d(rp[r]);
            res[r] = (16 * K + 4 * q + r < f) ? (float)((double)br[r] - y) : 0.f;
        }
        rc[K] = row_to_col(res);
    }
}

template <int T, int E>
__device__ __forceinline__ float stage_at(const char *stage, int k, int j) {
    using G = Geo<T, E>;
    if constexpr (E == PREC_FP32)
        return reinterpret_cast<const float *>(stage)[k * G::YR + j];
    else
        return from16<E>(reinterpret_cast<const uint16_t *>(stage)[G::yt_row(j) * G::YT_H + k]);
}

// r_j += sum_k y_k[j] e_k over the staged chunk, for factor j.
template <int T, int E>
__device__ __forceinline__ double stage_row_dot(const char *stage, int j, const double *eb) {
    using G = Geo<T, E>;
    double acc = 0.0;
    if constexpr (E == PREC_FP32) {
        const float *Ys = reinterpret_cast<const float *>(stage);
#pragma unroll 8
        for (int k = 0; k < CHUNK; ++k) acc += (double)Ys[k * G::YR + j] * eb[k];
    } else {
        const u32x4 *p = reinterpret_cast<const u32x4 *>(reinterpret_cast<const uint16_t *>(stage) + G::yt_row(j) * G::YT_H);
#pragma unroll
        for (int v = 0; v < CHUNK / 8; ++v) {
            const u32x4 w = p[v];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                acc += (double)from16<E>((uint16_t)(w[e] & 0xffffu)) * eb[8 * v + 2 * e];
                acc += (double)from16<E>((uint16_t)(w[e] >> 16)) * eb[8 * v + 2 * e + 1];
            }
        }
    }
    return acc;
}

template <int T, int J, int NJ>
__device__ __forceinline__ void lanes_to_colform(const double (&acc)[NJ], int f, float reg, const float (&xc)[T],
                                                 float (&rc)[T]) {
    if constexpr (J < T) {
        // r[16J + c] sits in register J/4 of lane 16 (J % 4) + c
        const double v = rows_bcast_d<(J & 3)>(acc[J >> 2]);
        const int i = 16 * J + (lane() & 15);
        rc[J] = i < f ? (float)(v - (double)reg * (double)xc[J]) : 0.f;
        lanes_to_colform<T, J + 1>(acc, f, reg, xc, rc);
    }
}

// Residual of one whole row without A: r = sum_i y_i (r_i - y_i . x) - reg x,
// the row's ratings restaged 32 at a time in precision E (the refinement
// restages the FP32 master rows, E = PREC_FP32, whatever the Gram's
// precision), reg the FP32 lambda * n the diagonal got. Products and sums in
// FP64. x column form; r column form (zero past f).
template <int T, int E>
__device__ __forceinline__ void residual_mf(const UpdateArgs &a, int64_t rs, int n, float reg, const float (&xc)[T],
                                            float (&rc)[T], char *lds) {
    using G = Geo<T, E>;
    constexpr int NJ = (G::FP + WAVE - 1) / WAVE;
    constexpr int HALF = G::FP / 2;
    const int l = lane(), f = a.f;
    const float *rbuf = reinterpret_cast<const float *>(lds) + 32;
    float *xs = reinterpret_cast<float *>(lds) + 64;
    double *eb = reinterpret_cast<double *>(lds + (64 + G::FP) * 4);
    const char *stage = lds + G::MISC_FLOATS * 4;
    if (l < 16) {
#pragma unroll
        for (int J = 0; J < T; ++J) xs[16 * J + l] = xc[J];
    }
    // (the first barrier of stage_chunk orders these writes before the reads)
    const int k = l & (CHUNK - 1), jb = (l >> 5) * HALF, je = min(f, jb + HALF);
    double acc[NJ];
#pragma unroll
    for (int m = 0; m < NJ; ++m) acc[m] = 0.0;
    for (int k0 = 0; k0 < n; k0 += CHUNK) {
        const int nv = min(CHUNK, n - k0);
        stage_chunk<T, E>(a, rs, k0, nv, lds);
        // e_k = r_k - y_k . x: lane (k, half) sums the factors [jb, je)
        double dot = 0.0;
        for (int j = jb; j < je; ++j) dot += (double)stage_at<T, E>(stage, k, j) * (double)xs[j];
        dot = halves_sum_d(dot);
        if (l < CHUNK) {
            float rk = rbuf[k];
            if constexpr (E != PREC_FP32) rk = from16<E>(to16<E>(rk));
            eb[k] = k < nv ? (double)rk - dot : 0.0;
        }
        wave_sync();
#pragma unroll
        for (int m = 0; m < NJ; ++m)
            if (WAVE * m + l < f) acc[m] += stage_row_dot<T, E>(stage, WAVE * m + l, eb);
        wave_sync();
    }
    lanes_to_colform<T, 0>(acc, f, reg, xc, rc);
}

// x += A^-1 r on the factor in the registers.
template <int T, int NT>
__device__ __forceinline__ void correct(const f32x4 (&a)[NT], int f, const float (&rc)[T], float (&xc)[T]) {
    f32x4 wr[T];
    fwd_solve<T>(a, f, rc, wr);
    float dc[T];
    back_solve<T>(a, f, wr, dc);
#pragma unroll
    for (int K = 0; K < T; ++K) xc[K] += dc[K];
}

__device__ __forceinline__ float colform_dot(const float *a, const float *b, int T) {
    float s = 0.f;
    for (int J = 0; J < T; ++J) s += a[J] * b[J];
    return row16_sum(s);
}

// ------------------------------------------------------------ mat-vec
// y = A v for the symmetric matrix held as tiles (layout above, before any
// factorisation), restricted to the leading f x f block. v and y in column
// form (zero past f).
//   diagonal (I,I): y[16I+4q+r] += A v[16I+c]        row-form partial
//   off (I,J):      y[16I+c]    += A v[16J+4q+r]     column-form partial
//                   y[16J+4q+r] += A v[16I+c]        row-form partial
// Block column J outermost: block J's row-form partial (its diagonal tile,
// then the tiles below it) and the column-form partials of the blocks below
// are complete at the end of step J, so y[J] is formed there and only one
// row-form partial and one row form of v are live at a time. Each partial
// sums its terms in ascending block order.
template <int T, int NT>
__device__ __forceinline__ void matvec_f32(const f32x4 (&A)[NT], int f, const float (&v)[T], float (&y)[T]) {
    const int c = lane() & 15;
    float pc[T];
#pragma unroll
    for (int I = 0; I < T; ++I) pc[I] = 0.f;
#pragma unroll
    for (int J = 0; J < T; ++J) {
        const f32x4 vr = col_to_row(v[J]);
        f32x4 pr = f32x4{0.f, 0.f, 0.f, 0.f};
        pr += A[tile_id(J, J)] * v[J];
#pragma unroll
        for (int I = J + 1; I < T; ++I) {
            const f32x4 t = A[tile_id(I, J)];
            pc[I] += t[0] * vr[0] + t[1] * vr[1] + t[2] * vr[2] + t[3] * vr[3];
            pr += t * v[I];
        }
        f32x4 s;
#pragma unroll
        for (int r = 0; r < 4; ++r) s[r] = row16_sum(pr[r]);
        const float yc = row_to_col(s) + q4_sum(pc[J]);
        y[J] = (16 * J + c < f) ? yc : 0.f;
    }
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
        // r.r == 0 (converged, or its square underflowed): the next beta
        // would be 0 / 0, so stop with the current x
        if (rsnew < tol || rsnew == 0.f) break;
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
// Next batch index from a device work queue ([0] next batch), uniform.
__device__ __forceinline__ int pop_batch(int *queue) {
    int b = 0;
    if (lane() == 0) b = atomicAdd(queue, 1);
    return __builtin_amdgcn_readfirstlane(__shfl(b, 0, WAVE));
}

// Exit protocol: the last wave out resets the queue ([1] counts waves that
// left) for the next launch, so a replayed graph needs no memset.
__device__ __forceinline__ void leave_queue(int *queue) {
    if (lane() == 0) {
        __threadfence();
        const int done = atomicAdd(queue + 1, 1);
        if (done == (int)gridDim.x - 1) {
            atomicExch(queue, 0);
            atomicExch(queue + 1, 0);
        }
    }
}

// ------------------------------------------------------------ dual solve
// A row with n ratings, n < f, has the same solution in the n x n space of
// its ratings. With Y the n x f matrix of the rated rows (16-bit) and
// mu = lambda n, the push-through identity gives
//   x = (Y^T Y + mu I)^-1 Y^T r = Y^T (Y Y^T + mu I)^-1 r,
// so the kernel factors the augmented [[Y Y^T + mu I, r], [r^T, 0]] in
// TD = n / 16 + 1 tiles instead of T, with the same register Cholesky.

// Sums of eight values over the 16 lanes of a row, scattered: lane c gets
// the total of v[c >> 1] (lanes c and c ^ 1 hold the same total). Exchange
// partners c ^ 8 (row_ror:8), c ^ 7 (row_half_mirror) and c ^ 2 halve the
// values carried; c ^ 1 adds the last pair.
__device__ __forceinline__ float rows16_scatter8(const float (&v)[8]) {
    const int c = lane() & 15;
    const bool b3 = (c & 8) != 0, b2 = (c & 4) != 0, b1 = (c & 2) != 0;
    float u[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) u[i] = (b3 ? v[4 + i] : v[i]) + dpp_mov<0x128>(b3 ? v[i] : v[4 + i]);
    float t[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) t[i] = (b2 ? u[2 + i] : u[i]) + dpp_mov<0x141>(b2 ? u[i] : u[2 + i]);
    float z = (b1 ? t[1] : t[0]) + dpp_mov<0x4E>(b1 ? t[0] : t[1]);
    return z + dpp_mov<0xB1>(z);
}

// Sums of four values over the 16 lanes of a row, scattered: lane c gets
// the total of v[c & 3] (the FP32 form of rows16_scatter_d).
__device__ __forceinline__ float rows16_scatter4(const float (&v)[4]) {
    const int c = lane() & 15;
    const bool b1 = (c & 2) != 0, b0 = (c & 1) != 0;
    const float u0 = (b1 ? v[2] : v[0]) + dpp_mov<0x4E>(b1 ? v[0] : v[2]);
    const float u1 = (b1 ? v[3] : v[1]) + dpp_mov<0x4E>(b1 ? v[1] : v[3]);
    float t = (b0 ? u1 : u0) + dpp_mov<0xB1>(b0 ? u0 : u1);
    t += dpp_mov<0x124>(t);
    t += dpp_mov<0x128>(t);
    return t;
}

// Operands of the dual Gram, one step of the factor dimension: lane (q, c)
// holds the 16-byte segment 4s + q of rated row 16I + c (zero past the
// ratings and past the segments that carry data). 16-bit rows: 8 factors,
// one 16x16x32 MFMA per tile. FP32 rows: 4 factors, four 16x16x4 MFMAs,
// the j-th taking factor 4 (4s + q) + j as its k = q (the contraction over
// the factors runs in any order, the same for both operands).
template <int TD, int T, int E>
__device__ __forceinline__ void dual_operands(const UpdateArgs &a, const int (&id)[TD], int s, int nseg,
                                              u32x4 (&F)[TD]) {
    using G = Geo<T, E>;
    const int seg = 4 * s + (lane() >> 4);
    constexpr int SEGW = E == PREC_FP32 ? 4 : 8;  // elements per 16-byte segment
#pragma unroll
    for (int I = 0; I < TD; ++I) {
        F[I] = u32x4{0u, 0u, 0u, 0u};
        if (id[I] >= 0 && seg < nseg) {
            if constexpr (E == PREC_FP32)
                F[I] = *reinterpret_cast<const u32x4 *>(a.Y + (int64_t)id[I] * G::FP + SEGW * seg);
            else
                F[I] = *reinterpret_cast<const u32x4 *>(a.Ys + (int64_t)id[I] * G::FP + SEGW * seg);
        }
    }
}

// One row through the dual system: x (column form of the T-tile layout,
// zero past f), a plain FP32 Cholesky solve of the 16-bit-formed (or FP32)
// dual matrix. Its matrix is at least lambda n I and its condition number at
// most 1 + |y|^2 / lambda over the rated rows, so the FP32 solve's error
// stays far below the rounding of the 16-bit Gram that every row of the
// mixed path carries; no refinement step is taken here.
// xs: 16T floats of this wave's LDS.
template <int TD, int T, int E>
__device__ __forceinline__ void dual_row(const UpdateArgs &a, int64_t rs, int n, float reg, float (&xc)[T],
                                         float *xs) {
    constexpr int NTD = TD * (TD + 1) / 2;
    int l = lane();
    // Opaque per row: lane masks of this path are formed here, not hoisted
    // out of the row loop where they would stay live across the other path.
    asm volatile("" : "+v"(l));
    const int q = l >> 4, c = l & 15;
    const int f = a.f;
    constexpr int SEGW = E == PREC_FP32 ? 4 : 8;
    const int nseg = (f + SEGW - 1) / SEGW, ks = (nseg + 3) >> 2;
    int id[TD];
    float rc[TD];  // ratings, column form, rounded to the Gram precision as the augmented Gram rounds them
#pragma unroll
    for (int I = 0; I < TD; ++I) {
        const int k = 16 * I + c;
        id[I] = k < n ? a.idx[rs + k] : -1;
        float v = k < n ? a.val[rs + k] : 0.f;
        if constexpr (E != PREC_FP32) v = from16<E>(to16<E>(v));
        rc[I] = v;
    }
    // K = Y Y^T: off-diagonal tiles come out row by lane (operands swapped)
    f32x4 acc[NTD];
#pragma unroll
    for (int t = 0; t < NTD; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int s = 0; s < ks; ++s) {
        u32x4 F[TD];
        dual_operands<TD, T, E>(a, id, s, nseg, F);
#pragma unroll
        for (int I = 0; I < TD; ++I)
#pragma unroll
            for (int J = 0; J <= I; ++J) {
                if constexpr (E == PREC_FP32) {
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        acc[tile_id(I, J)] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                            __uint_as_float(F[J][j]), __uint_as_float(F[I][j]), acc[tile_id(I, J)], 0, 0, 0);
                } else if constexpr (E == PREC_BF16) {
                    acc[tile_id(I, J)] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(bf16x8, F[J]), __builtin_bit_cast(bf16x8, F[I]), acc[tile_id(I, J)], 0, 0, 0);
                } else {
                    acc[tile_id(I, J)] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                        __builtin_bit_cast(f16x8, F[J]), __builtin_bit_cast(f16x8, F[I]), acc[tile_id(I, J)], 0, 0, 0);
                }
            }
    }
    add_diag<TD>(acc, n, reg);
    // row n <- r: row by lane in the tiles left of the diagonal (lanes
    // c == n % 16), row and column n in the last diagonal tile
    {
        const int nr = n & 15;
#pragma unroll
        for (int J = 0; J + 1 < TD; ++J) {
            const f32x4 rr = col_to_row(rc[J]);
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[tile_id(TD - 1, J)][r] = (c == nr) ? rr[r] : acc[tile_id(TD - 1, J)][r];
        }
        const f32x4 rr = col_to_row(rc[TD - 1]);
        constexpr int D = tile_id(TD - 1, TD - 1);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float v = (c == nr) ? rr[r] : acc[D][r];  // column n
            v = (4 * q + r == nr) ? rc[TD - 1] : v;   // row n
            acc[D][r] = v;
        }
    }
    f32x4 zr[TD];
    chol_regs<TD, NTD, (E != PREC_FP32)>(acc, n, zr);
    float al[TD];
    back_solve<TD>(acc, n, zr, al);
    // x = Y^T alpha: lane (q, c) multiplies segment 4s + q of row 16I + c by
    // alpha[16I + c]; after the scattered sum over c, lane c holds factor
    // 8 (4s + q) + c / 2 (16-bit) or 4 (4s + q) + c % 4 (FP32).
    for (int s = 0; s < ks; ++s) {
        u32x4 F[TD];
        dual_operands<TD, T, E>(a, id, s, nseg, F);
        const int seg = 4 * s + q;
        if constexpr (E == PREC_FP32) {
            float p[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int I = 0; I < TD; ++I)
#pragma unroll
                for (int j = 0; j < 4; ++j) p[j] = fmaf(__uint_as_float(F[I][j]), al[I], p[j]);
            const float z = rows16_scatter4(p);
            if (c < 4 && seg < nseg) xs[4 * seg + c] = z;
        } else {
            float p[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int I = 0; I < TD; ++I)
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    p[2 * w] = fmaf(from16<E>((uint16_t)(F[I][w] & 0xffffu)), al[I], p[2 * w]);
                    p[2 * w + 1] = fmaf(from16<E>((uint16_t)(F[I][w] >> 16)), al[I], p[2 * w + 1]);
                }
            const float z = rows16_scatter8(p);
            if ((c & 1) == 0 && seg < nseg) xs[8 * seg + (c >> 1)] = z;
        }
    }
    wave_sync();
#pragma unroll
    for (int J = 0; J < T; ++J) xc[J] = (16 * J + c < f) ? xs[16 * J + c] : 0.f;
    wave_sync();  // x read before the next row's use of xs
}

// The dual system of the fewest tiles that hold n + 1 rows.
template <int TD, int TDMAX, int T, int E>
__device__ __forceinline__ void dual_pick(const UpdateArgs &a, int64_t rs, int n, float reg, float (&xc)[T],
                                          float *xs) {
    if constexpr (TD < TDMAX) {
        if (n < 16 * TD) dual_row<TD, T, E>(a, rs, n, reg, xc, xs);
        else dual_pick<TD + 1, TDMAX, T, E>(a, rs, n, reg, xc, xs);
    } else {
        dual_row<TD, T, E>(a, rs, n, reg, xc, xs);
    }
}

// Every work unit of queue batch b of one model.
template <int T, int E, int S>
__device__ __forceinline__ void run_batch(const UpdateArgs &a, int b, char *lds) {
    using G = Geo<T, E>;
    const int l = lane();
    const int f = a.f;
    float *Ascr = a.scratch ? a.scratch + (int64_t)blockIdx.x * G::NT * 256 : nullptr;
    f32x4 acc[G::NT];
    {
        const int u0 = a.batch_begin[b], u1 = a.batch_begin[b + 1];
        for (int u = u0; u < u1; ++u) {
            const int4 w = a.units[u];
            const int row = w.x;
            const int n = a.ptr[row + 1] - a.ptr[row];  // whole row (lambda * n_u)
            float x[T];
#pragma unroll
            for (int J = 0; J < T; ++J) x[J] = 0.f;
            // Warm-started refinement of a whole row: the row's previous
            // value x0 (its FP32 row of X, 16T floats) is copied to the LDS
            // vector area (xs) by an LDS-DMA load issued before the Gram, so
            // its memory latency runs under the Gram; lane l copies floats
            // 4l .. 4l+3. Only this wave writes that area, and it waits for
            // the copy (vmcnt) before reading it, in this same iteration.
            // (Parts of split rows, which may leave the iteration before the
            // solve, load x0 from X when they solve.)
            constexpr bool WARM_KERNEL = S == SOLVE_CHOL_WARM || S == SOLVE_CHOL_REFINE;
            const bool warm_row =
                WARM_KERNEL && (S == SOLVE_CHOL_WARM || (a.refine_warm && a.refine > 0)) && n > 0 && w.w < 0;
            float *xs = reinterpret_cast<float *>(lds) + 64;
            if constexpr (WARM_KERNEL) {
                if (warm_row && l < G::FP / 4)
                    __builtin_amdgcn_global_load_lds(a.X + (int64_t)row * G::FP + 4 * l,
                                                     (__attribute__((address_space(3))) void *)xs, 16, 0, 0);
            }
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
                        // T tiles' loads in flight before their adds (one
                        // register reused for every tile would wait for each
                        // load in turn)
#pragma unroll
                        for (int t0 = 0; t0 < G::NT; t0 += T) {
                            f32x4 part_t[T];
#pragma unroll
                            for (int k = 0; k < T; ++k)
                                if (t0 + k < G::NT)
                                    part_t[k] = *reinterpret_cast<const f32x4 *>(src + (t0 + k) * 256 + 4 * l);
#pragma unroll
                            for (int k = 0; k < T; ++k)
                                if (t0 + k < G::NT) acc[t0 + k] += part_t[k];
                        }
                    }
                    if (l == 0) atomicExch(a.counters + w.w, 0);
                }
                const float lam = a.lambda_rows ? a.lambda_rows[row] : a.lambda;
                const float reg = lam * (float)n;
                add_diag<T>(acc, f, reg);
                bool ok = true;
                if constexpr (S == SOLVE_CHOL || S == SOLVE_CHOL_REFINE || S == SOLVE_CHOL_WARM) {
                    // Refinement residual: restage the ratings of a whole row
                    // up to restage_max long, else keep a copy of the tiles.
                    const bool restage = w.w < 0 && n <= a.restage_max;
                    // Warm-started first step: the FP64 residual of the row's
                    // previous value x0 (still in X) is formed from the tiles
                    // before they are factorised, and x = x0 + A^-1 r0.
                    // The augmented row then carries b - A x0 instead of b, so the
                    // factorisation itself yields L^-1 (b - A x0).
                    const bool warm =
                        S == SOLVE_CHOL_WARM || (S == SOLVE_CHOL_REFINE && a.refine_warm && a.refine > 0);
                    float x0[T];
                    if constexpr (S == SOLVE_CHOL_REFINE) {
                        if (!restage && a.refine > (warm ? 1 : 0)) {
#pragma unroll
                            for (int t = 0; t < G::NT; ++t)
                                *reinterpret_cast<f32x4 *>(Ascr + t * 256 + 4 * l) = acc[t];
                        }
                    }
                    if constexpr (S == SOLVE_CHOL_REFINE || S == SOLVE_CHOL_WARM) {
                        if (warm) {
                            if (warm_row) {
                                // the copy of x0 has landed in xs (and no other
                                // load is outstanding after a whole row's Gram)
                                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                                load_colform<T>(xs, f, x0);
                            } else {
                                load_colform<T>(a.X + (int64_t)row * G::FP, f, x0);
                            }
                            warm_row_f<T>(acc, f, x0, xs, reinterpret_cast<float *>(lds + G::MISC_FLOATS * 4));
                        }
                    }
                    f32x4 zr[T];
                    // the 16-bit-Gram paths factor with split-FP16 trailing updates
                    ok = chol_regs<T, G::NT, (E != PREC_FP32)>(acc, f, zr);
                    back_solve<T>(acc, f, zr, x);
                    if (warm) {
                        // x0 again (not kept live across the factorisation): from
                        // xs, where the residual wrote the same values, or from X
                        if (warm_row) load_colform<T>(xs, f, x0);
                        else load_colform<T>(a.X + (int64_t)row * G::FP, f, x0);
#pragma unroll
                        for (int J = 0; J < T; ++J) x[J] += x0[J];
                    }
                    if constexpr (S == SOLVE_CHOL_REFINE) {
                        for (int s = warm ? 1 : 0; s < a.refine; ++s) {
                            float rc[T];
                            // restaged from the FP32 master rows: the residual of the
                            // unrounded system, which the correction solves then
                            // converge to (the 16-bit Gram's factor preconditions them)
                            if (restage) residual_mf<T, PREC_FP32>(a, w.y, n, reg, x, rc, lds);
                            else residual_copy<T>(Ascr, f, x, rc);
                            correct<T>(acc, f, rc, x);
                        }
                    }
                    if (a.Lsave) {
                        float *dst = a.Lsave + (int64_t)row * G::NT * 256;
#pragma unroll
                        for (int t = 0; t < G::NT; ++t) *reinterpret_cast<f32x4 *>(dst + t * 256 + 4 * l) = acc[t];
                    }
                } else {
                    float bvec[T];
                    extract_b<T>(acc, f, bvec);
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
}

template <int T, int E, int S>
__device__ __forceinline__ void update_body(const UpdateArgs &a) {
    extern __shared__ __attribute__((aligned(16))) char lds[];
    for (;;) {
        const int b = pop_batch(a.queue);
        if (b >= a.nbatches) break;
        run_batch<T, E, S>(a, b, lds);
    }
    leave_queue(a.queue);
}

template <int T, int E, int S>
__global__ __launch_bounds__(WAVE) void update_kernel(UpdateArgs a) {
    update_body<T, E, S>(a);
}

// Light rows through the dual system: whole rows with fewer than
// 16 * DUAL_TD ratings (and empty rows, x = 0), the units of their own
// schedule, launched beside the update kernel of the same half-sweep. It
// carries none of the primal Gram and solve code, and the update kernel
// none of its code. Four waves per SIMD (128 VGPRs; at T = 8 and 9 one or
// two registers spill): its row is mostly VALU work on few live registers,
// and four waves run it faster than three (users half -1.8% at f = 100).
template <int T, int E, int S>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(4, 4))) void dual_kernel(UpdateArgs a) {
    using G = Geo<T, E>;
    extern __shared__ __attribute__((aligned(16))) char lds[];
    const int l = lane();
    float *xs = reinterpret_cast<float *>(lds) + 64;
    for (;;) {
        const int b = pop_batch(a.queue);
        if (b >= a.nbatches) break;
        const int u0 = a.batch_begin[b], u1 = a.batch_begin[b + 1];
        for (int u = u0; u < u1; ++u) {
            const int row = a.units[u].x;
            const int n = a.ptr[row + 1] - a.ptr[row];
            float x[T];
#pragma unroll
            for (int J = 0; J < T; ++J) x[J] = 0.f;
            if (n > 0) {
                const float reg = (a.lambda_rows ? a.lambda_rows[row] : a.lambda) * (float)n;
                dual_pick<1, G::DUAL_TD, T, E>(a, a.ptr[row], n, reg, x, xs);
                bool finite = true;
#pragma unroll
                for (int J = 0; J < T; ++J) finite = finite && isfinite(x[J]);
                if (__any(!finite)) {
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
    leave_queue(a.queue);
}

// The refinement variant holds the factor, the solution and the FP64
// residual pieces at once; at two waves per SIMD it trades a small spill
// for twice the latency hiding of the one-wave form.
template <int T, int E, int S>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(2, 2))) void update_kernel_w2(UpdateArgs a) {
    update_body<T, E, S>(a);
}

}  // namespace als
