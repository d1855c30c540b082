// IC-GN subset solver for gfx950: one 64-lane wavefront per subset.
//
// Reference semantics: src/math/optimization_engine.cpp (calculate_deformation
// :55-118, estimate_initial_guess :120-179, sample_warped_subset :191-269,
// solve_icgn :272-465, evaluate_znssd :467-528, solve_simplex :531-651) and the
// fused reductions of include/semper/simd.hpp. Subset preparation
// (precompute_subset_fast, src/math/subset_precomputer.cpp:159-265) is the
// kernel prologue.
//
// Layout: lane l owns subset pixels i = l + 64 j (j < PPL), row-major over the
// subset, and keeps their normalised reference value and scaled gradients in
// registers for the whole solve. Each step warps, guards and interpolates the
// lane's pixels (the deformed image is read through L1/L2: its bytes are shared
// by ~(d/step)^2 neighbouring subsets) and forms the ZNSSD statistics with
// butterfly reductions. The 6x6 Gauss-Newton step and the warp composition run
// redundantly in every lane on wave-uniform values, so control flow never
// diverges.
//
// IC-GN iterations and Nelder-Mead evaluations share one sampling site: the
// solve is a uniform state machine (IC-GN run, Simplex, IC-GN rescue run) that
// hands one warp per step to a single evaluator. The rarely used state (the
// Simplex vertices, the 6x6 inverse of a masked iteration) lives in LDS.
//
// Arithmetic follows the reference operation by operation (multiply then add,
// no contraction; the warp position summed in double), so a sample and its
// interpolated value equal the CPU restatement's; only the order of the
// reductions differs.
#pragma once
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "solver.hpp"
#include "wave.hpp"

#pragma clang fp contract(off)

// pixels per lane whose Keys 4x4 gathers the fast path issues together: over
// the FP32 plane (16 dwords a pixel), and over the u8 or coefficient plane
// (4 or 8 dwords a pixel)
#ifndef SDIC_K4_BATCH_F32
#define SDIC_K4_BATCH_F32 2
#endif
#ifndef SDIC_K4_BATCH_NARROW
#define SDIC_K4_BATCH_NARROW 4
#endif

namespace sdic {

namespace {

// Keys a = -0.5 weights (image_processor.cpp:188-195)
__device__ __forceinline__ void keys4(float s, float* w) {
    const float s2 = s * s;
    const float s3 = s2 * s;
    w[0] = -0.5f * s3 + s2 - 0.5f * s;
    w[1] = 1.5f * s3 - 2.5f * s2 + 1.0f;
    w[2] = -1.5f * s3 + 2.0f * s2 + 0.5f * s;
    w[3] = 0.5f * s3 - 0.5f * s2;
}
// Keys 4th-order 6x6 kernel pieces (image_processor.cpp:71-82)
__device__ __forceinline__ float k6f0(float s) {
    return (4.0f / 3.0f) * s * s * s - (7.0f / 3.0f) * s * s + 1.0f;
}
__device__ __forceinline__ float k6f1(float s) {
    return -(7.0f / 12.0f) * s * s * s + 3.0f * s * s - (59.0f / 12.0f) * s + 2.5f;
}
__device__ __forceinline__ float k6f2(float s) {
    return (1.0f / 12.0f) * s * s * s - (2.0f / 3.0f) * s * s + (7.0f / 4.0f) * s - 1.5f;
}

// Interpolation inside the 4 px guard, where neither kernel demotes to
// bilinear (the guard keeps x in [3.5, w - 4.5)).
template <bool K6, typename TD>
__device__ __forceinline__ float interp(const TD* __restrict__ img, int w, float x, float y) {
    const int xi = (int)x, yi = (int)y;
    const float dx = x - (float)xi, dy = y - (float)yi;
    if constexpr (!K6) {
        float wx[4], wy[4];
        keys4(dx, wx);
        keys4(dy, wy);
        float val = 0.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const TD* r = img + (yi + j - 1) * w + xi;
            float rv = 0.0f;
            rv += (float)r[-1] * wx[0];
            rv += (float)r[0] * wx[1];
            rv += (float)r[1] * wx[2];
            rv += (float)r[2] * wx[3];
            val += rv * wy[j];
        }
        return val;
    } else {
        const float wx[6] = {k6f2(dx + 2.0f), k6f1(dx + 1.0f), k6f0(dx),
                             k6f0(1.0f - dx), k6f1(2.0f - dx), k6f2(3.0f - dx)};
        const float wy[6] = {k6f2(dy + 2.0f), k6f1(dy + 1.0f), k6f0(dy),
                             k6f0(1.0f - dy), k6f1(2.0f - dy), k6f2(3.0f - dy)};
        float val = 0.0f;
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const TD* r = img + (yi + j - 2) * w + xi;
            float rv = 0.0f;
            rv += (float)r[-2] * wx[0];
            rv += (float)r[-1] * wx[1];
            rv += (float)r[0] * wx[2];
            rv += (float)r[1] * wx[3];
            rv += (float)r[2] * wx[4];
            rv += (float)r[3] * wx[5];
            val += rv * wy[j];
        }
        return val;
    }
}

__device__ __forceinline__ bool inside_guard(float x, float y, int w, int h) {
    const float rx = x + 0.5f, ry = y + 0.5f;
    return rx >= 4.0f && rx < (float)(w - 4) && ry >= 4.0f && ry < (float)(h - 4);
}

// optimization_engine.cpp:19-21
__device__ __forceinline__ float warped(int c, float a, float b, float t, float x, float y) {
    return (float)((double)c + ((double)a * (double)x + (double)b * (double)y + (double)t));
}

// ---- the lean path's arithmetic ------------------------------------------
// The Keys weights in Horner form and the stencil as fused multiply-adds: the
// same polynomials as keys4 / k6f*, rounded once per step instead of twice.
__device__ __forceinline__ void keys4_fma(float s, float* w) {
    const float s2 = s * s;
    w[0] = s * __builtin_fmaf(__builtin_fmaf(-0.5f, s, 1.0f), s, -0.5f);
    w[1] = __builtin_fmaf(s2, __builtin_fmaf(1.5f, s, -2.5f), 1.0f);
    w[2] = s * __builtin_fmaf(__builtin_fmaf(-1.5f, s, 2.0f), s, 0.5f);
    w[3] = s2 * __builtin_fmaf(0.5f, s, -0.5f);
}
__device__ __forceinline__ float k6f0_fma(float s) {
    return __builtin_fmaf(s * s, __builtin_fmaf(4.0f / 3.0f, s, -7.0f / 3.0f), 1.0f);
}
__device__ __forceinline__ float k6f1_fma(float s) {
    return __builtin_fmaf(s, __builtin_fmaf(s, __builtin_fmaf(-7.0f / 12.0f, s, 3.0f), -59.0f / 12.0f), 2.5f);
}
__device__ __forceinline__ float k6f2_fma(float s) {
    return __builtin_fmaf(s, __builtin_fmaf(s, __builtin_fmaf(1.0f / 12.0f, s, -2.0f / 3.0f), 7.0f / 4.0f), -1.5f);
}

template <bool K6, typename TD>
__device__ __forceinline__ float interp_fma(const TD* __restrict__ img, int w, float x, float y) {
    const int xi = (int)x, yi = (int)y;
    const float dx = x - (float)xi, dy = y - (float)yi;
    if constexpr (!K6) {
        float wx[4], wy[4];
        keys4_fma(dx, wx);
        keys4_fma(dy, wy);
        float val = 0.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const TD* r = img + (yi + j - 1) * w + xi;
            float rv = (float)r[-1] * wx[0];
            rv = __builtin_fmaf((float)r[0], wx[1], rv);
            rv = __builtin_fmaf((float)r[1], wx[2], rv);
            rv = __builtin_fmaf((float)r[2], wx[3], rv);
            val = __builtin_fmaf(rv, wy[j], val);
        }
        return val;
    } else {
        const float wx[6] = {k6f2_fma(dx + 2.0f), k6f1_fma(dx + 1.0f), k6f0_fma(dx),
                             k6f0_fma(1.0f - dx), k6f1_fma(2.0f - dx), k6f2_fma(3.0f - dx)};
        const float wy[6] = {k6f2_fma(dy + 2.0f), k6f1_fma(dy + 1.0f), k6f0_fma(dy),
                             k6f0_fma(1.0f - dy), k6f1_fma(2.0f - dy), k6f2_fma(3.0f - dy)};
        float val = 0.0f;
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const TD* r = img + (yi + j - 2) * w + xi;
            float rv = (float)r[-2] * wx[0];
            rv = __builtin_fmaf((float)r[-1], wx[1], rv);
            rv = __builtin_fmaf((float)r[0], wx[2], rv);
            rv = __builtin_fmaf((float)r[1], wx[3], rv);
            rv = __builtin_fmaf((float)r[2], wx[4], rv);
            rv = __builtin_fmaf((float)r[3], wx[5], rv);
            val = __builtin_fmaf(rv, wy[j], val);
        }
        return val;
    }
}

// The same interpolation through a buffer resource over the deformed plane:
// one 32-bit byte offset per sample (yi - 1) * w + xi - 1 in a VGPR, the row
// step j * w in the wave-uniform SGPR offset, and one load per stencil row
// (4 bytes / 4 floats for Keys 4x4, 8 bytes / 6 floats for Keys 6x6). The
// sample lies inside the 4 px guard, so every row read is inside the plane.
__device__ __forceinline__ float ub(uint32_t v, int k) { return (float)((v >> (8 * k)) & 0xFFu); }

template <bool K6, typename TD>
__device__ __forceinline__ float interp_buf(__amdgpu_buffer_rsrc_t rs, int w, float x, float y) {
    const int xi = (int)x, yi = (int)y;
    const float dx = __builtin_amdgcn_fractf(x), dy = __builtin_amdgcn_fractf(y);
    constexpr int R = K6 ? 6 : 4, H = K6 ? 2 : 1;
    const int base = ((int)__umul24((unsigned)(yi - H), (unsigned)w) + xi - H) * (int)sizeof(TD);
    float wx[R], wy[R];
    if constexpr (!K6) {
        keys4_fma(dx, wx);
        keys4_fma(dy, wy);
    } else {
        const float ax[6] = {k6f2_fma(dx + 2.0f), k6f1_fma(dx + 1.0f), k6f0_fma(dx),
                             k6f0_fma(1.0f - dx), k6f1_fma(2.0f - dx), k6f2_fma(3.0f - dx)};
        const float ay[6] = {k6f2_fma(dy + 2.0f), k6f1_fma(dy + 1.0f), k6f0_fma(dy),
                             k6f0_fma(1.0f - dy), k6f1_fma(2.0f - dy), k6f2_fma(3.0f - dy)};
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            wx[k] = ax[k];
            wy[k] = ay[k];
        }
    }
    float val = 0.0f;
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int so = j * w * (int)sizeof(TD);
        float px[R];
        if constexpr (sizeof(TD) == 1) {
            if constexpr (!K6) {
                const uint32_t b = __builtin_amdgcn_raw_buffer_load_b32(rs, base, so, 0);
#pragma unroll
                for (int k = 0; k < 4; ++k) px[k] = ub(b, k);
            } else {
                const auto b = __builtin_amdgcn_raw_buffer_load_b64(rs, base, so, 0);
#pragma unroll
                for (int k = 0; k < 4; ++k) px[k] = ub((uint32_t)b[0], k);
                px[4] = ub((uint32_t)b[1], 0);
                px[5] = ub((uint32_t)b[1], 1);
            }
        } else {
            const auto b = __builtin_amdgcn_raw_buffer_load_b128(rs, base, so, 0);
#pragma unroll
            for (int k = 0; k < 4; ++k) px[k] = __uint_as_float((uint32_t)b[k]);
            if constexpr (K6) {
                const auto c = __builtin_amdgcn_raw_buffer_load_b64(rs, base + 16, so, 0);
                px[4] = __uint_as_float((uint32_t)c[0]);
                px[5] = __uint_as_float((uint32_t)c[1]);
            }
        }
        float rv = px[0] * wx[0];
#pragma unroll
        for (int k = 1; k < R; ++k) rv = __builtin_fmaf(px[k], wx[k], rv);
        val = __builtin_fmaf(rv, wy[j], val);
    }
    return val;
}

// Keys 4x4 from the row-cubic coefficient plane (per pixel the cubics in s of
// the cell to its right on its row and the next, four fp16 each, exact for
// u8 taps): two 16-byte loads per sample (rows yi - 1, yi and yi + 1, yi + 2),
// each row's value by Horner in dx, the four rows combined with the Keys
// weights in dy. The same interpolant as interp_buf; the roundings differ
// (Horner against the tap-weight sum).
typedef _Float16 sdic_h4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float coef_row(uint32_t lo, uint32_t hi, float dx) {
    const sdic_h4 c = __builtin_bit_cast(sdic_h4, make_uint2(lo, hi));
    float rv = __builtin_fmaf((float)c.w, dx, (float)c.z);
    rv = __builtin_fmaf(rv, dx, (float)c.y);
    return __builtin_fmaf(rv, dx, (float)c.x);
}

__device__ __forceinline__ float interp_coef(__amdgpu_buffer_rsrc_t rs, int w, float x, float y) {
    const int xi = (int)x, yi = (int)y;
    const float dx = __builtin_amdgcn_fractf(x), dy = __builtin_amdgcn_fractf(y);
    const int base = ((int)__umul24((unsigned)(yi - 1), (unsigned)w) + xi) * 16;
    float wy[4];
    keys4_fma(dy, wy);
    const auto a = __builtin_amdgcn_raw_buffer_load_b128(rs, base, 0, 0);
    const auto b = __builtin_amdgcn_raw_buffer_load_b128(rs, base, 2 * w * 16, 0);
    float val = 0.0f;
    val = __builtin_fmaf(coef_row((uint32_t)a[0], (uint32_t)a[1], dx), wy[0], val);
    val = __builtin_fmaf(coef_row((uint32_t)a[2], (uint32_t)a[3], dx), wy[1], val);
    val = __builtin_fmaf(coef_row((uint32_t)b[0], (uint32_t)b[1], dx), wy[2], val);
    val = __builtin_fmaf(coef_row((uint32_t)b[2], (uint32_t)b[3], dx), wy[3], val);
    return val;
}

// (Row offsets use the full-rate 24-bit multiply: the row and the width are
// below 2^24 and their product, a pixel index, below 2^31.)
// Keys 4x4 in two phases, so that a wave keeps the gathers of several pixels
// in flight: k4_issue computes the cell and issues the stencil loads (the raw
// dwords: 1 per row from the u8 plane, 4 per row from the FP32 plane, two
// 16-byte row pairs from the coefficient plane), k4_eval combines them with
// exactly the operations of interp_buf / interp_coef (so the value is theirs
// bit for bit).
struct K4Rows {
    uint32_t r[4][4];
    float dx, dy;
};

template <typename TD, bool COEF>
__device__ __forceinline__ void k4_issue(__amdgpu_buffer_rsrc_t rs, int w, float x, float y, K4Rows& k) {
    const int xi = (int)x, yi = (int)y;
    k.dx = __builtin_amdgcn_fractf(x);
    k.dy = __builtin_amdgcn_fractf(y);
    if constexpr (COEF) {
        const int base = ((int)__umul24((unsigned)(yi - 1), (unsigned)w) + xi) * 16;
        const auto a = __builtin_amdgcn_raw_buffer_load_b128(rs, base, 0, 0);
        const auto b = __builtin_amdgcn_raw_buffer_load_b128(rs, base, 2 * w * 16, 0);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            k.r[0][c] = (uint32_t)a[c];
            k.r[1][c] = (uint32_t)b[c];
        }
    } else if constexpr (sizeof(TD) == 1) {
        const int base = (int)__umul24((unsigned)(yi - 1), (unsigned)w) + xi - 1;
#pragma unroll
        for (int j = 0; j < 4; ++j) k.r[j][0] = __builtin_amdgcn_raw_buffer_load_b32(rs, base, j * w, 0);
    } else {
        const int base = ((int)__umul24((unsigned)(yi - 1), (unsigned)w) + xi - 1) * 4;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const auto b = __builtin_amdgcn_raw_buffer_load_b128(rs, base, j * w * 4, 0);
#pragma unroll
            for (int c = 0; c < 4; ++c) k.r[j][c] = (uint32_t)b[c];
        }
    }
}

template <typename TD, bool COEF>
__device__ __forceinline__ float k4_eval(const K4Rows& k) {
    float wy[4];
    keys4_fma(k.dy, wy);
    float val = 0.0f;
    if constexpr (COEF) {
        // rows yi - 1, yi in k.r[0], rows yi + 1, yi + 2 in k.r[1]
#pragma unroll
        for (int j = 0; j < 4; ++j)
            val = __builtin_fmaf(coef_row(k.r[j >> 1][2 * (j & 1)], k.r[j >> 1][2 * (j & 1) + 1], k.dx), wy[j], val);
    } else {
        float wx[4];
        keys4_fma(k.dx, wx);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float px[4];
#pragma unroll
            for (int c = 0; c < 4; ++c)
                px[c] = sizeof(TD) == 1 ? ub(k.r[j][0], c) : __uint_as_float(k.r[j][c]);
            float rv = px[0] * wx[0];
#pragma unroll
            for (int c = 1; c < 4; ++c) rv = __builtin_fmaf(px[c], wx[c], rv);
            val = __builtin_fmaf(rv, wy[j], val);
        }
    }
    return val;
}

// interp<false>'s arithmetic (keys4 weights, products summed in stencil
// order, no fused multiply-adds) on rows loaded by k4_issue from a u8 or
// FP32 plane: its value bit for bit (the fraction fract(x) equals
// x - (int)x for the positive positions inside the guard).
template <typename TD>
__device__ __forceinline__ float k4_eval_ref(const K4Rows& k) {
    float wx[4], wy[4];
    keys4(k.dx, wx);
    keys4(k.dy, wy);
    float val = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        float px[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) px[c] = sizeof(TD) == 1 ? ub(k.r[j][0], c) : __uint_as_float(k.r[j][c]);
        float rv = 0.0f;
        rv += px[0] * wx[0];
        rv += px[1] * wx[1];
        rv += px[2] * wx[2];
        rv += px[3] * wx[3];
        val += rv * wy[j];
    }
    return val;
}

// The warp position of `warped` with the same roundings: a*x is exact in
// double (a 24-bit by 6-bit product), so fma(a, x, b*y) rounds exactly as
// a*x + b*y does.
__device__ __forceinline__ float warped_fma(double c, double a, double b, double t, double x, double y) {
    return (float)(c + (__builtin_fma(a, x, b * y) + t));
}

typedef _Float16 sdic_h2 __attribute__((ext_vector_type(2)));

struct Res {
    float p[6];
    float corr;
    int status;
    int iters;
};

struct WaveLds {
    float H0[36];     // the pool's step matrix
    float P[7 * 6];   // Simplex vertices (vertex k, coordinate j at 6 k + j)
};

// Sums over the NW waves that solve one subset together (the full solver):
// each wave's butterfly sum, then the waves' partials in wave order, so that
// every wave holds the same bits and takes the same branches. red is
// 2 x NW x kGroupSlots floats of LDS; par alternates its halves, so that one
// sum's writes cannot overtake a slower wave still reading the previous sum
// (the barrier between them orders every wave's reads before the next writes).
constexpr int kGroupSlots = 24;
constexpr int kSlowWaves = 4;   // waves per subset in the full solver

template <int NW, int K>
__device__ __forceinline__ void group_sum_n(float (&v)[K], float* red, int& par) {
    static_assert(K <= kGroupSlots, "too many sums at once");
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    if constexpr (NW > 1) {
        const int wid = (int)(threadIdx.x >> 6);
        float* b = red + par * NW * kGroupSlots;
        if ((threadIdx.x & 63) == 0) {
#pragma unroll
            for (int k = 0; k < K; ++k) b[wid * kGroupSlots + k] = v[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; ++k) {
            float acc = b[k];
#pragma unroll
            for (int w = 1; w < NW; ++w) acc += b[w * kGroupSlots + k];
            v[k] = acc;
        }
        par ^= 1;
    }
}

template <int NW>
__device__ __forceinline__ float group_sum(float v, float* red, int& par) {
    float a[1] = {v};
    group_sum_n<NW, 1>(a, red, par);
    return a[0];
}

template <int NW>
__device__ __forceinline__ int group_sum_i(int v, float* red, int& par) {
    v = wave_sum_i(v);
    if constexpr (NW > 1) {
        const int wid = (int)(threadIdx.x >> 6);
        float* b = red + par * NW * kGroupSlots;
        if ((threadIdx.x & 63) == 0) b[wid * kGroupSlots] = __int_as_float(v);
        __syncthreads();
        int acc = __float_as_int(b[0]);
#pragma unroll
        for (int w = 1; w < NW; ++w) acc += __float_as_int(b[w * kGroupSlots]);
        v = acc;
        par ^= 1;
    }
    return v;
}

enum Phase : int { PH_ICGN = 0, PH_SIMPLEX = 1, PH_DONE = 2 };
enum SStage : int { S_INIT = 0, S_REFLECT = 1, S_EXPAND = 2, S_CONTRACT = 3, S_SHRINK = 4 };

template <int PPL, bool K6, typename TD, int NW>
__device__ __forceinline__ void full_solve(const Task& t, const PoolEntry* __restrict__ pool,
                                          const float* __restrict__ refI,
                                          const float* __restrict__ refGx,
                                          const float* __restrict__ refGy,
                                          const uint16_t* __restrict__ gx16, const uint16_t* __restrict__ gy16, bool bf16,
                                          const double* __restrict__ chol, const TD* __restrict__ def,
                                          int w, int h, int dim, bool lm, float conv_tol,
                                          WaveLds& L, float* red, SolveOut& out) {
    const int lane = threadIdx.x & 63;
    // NW waves solve the subset together: thread tid holds pixels tid + T j
    constexpr int T = 64 * NW;
    constexpr int SL = PPL / NW;
    static_assert(PPL % NW == 0, "the pixel slots must split over the waves");
    const int tid = (int)threadIdx.x;
    int par = 0;   // the half of red the next group sum writes
    const PoolEntry* pe = pool + t.pool;
    out.init = 0;
    out.simplex = 0;
    if (!pe->valid) return;
    const int half = dim / 2;
    if (t.cx - half < 0 || t.cx + half >= w || t.cy - half < 0 || t.cy + half >= h) return;
    const int n = dim * dim;
    // the deformed plane as a buffer resource (the Keys 4x4 gathers)
    const __amdgpu_buffer_rsrc_t drs =
        __builtin_amdgcn_make_buffer_rsrc((void*)def, (short)0, (int)((size_t)w * h * sizeof(TD)), 0x00020000);
    // slots whose gathers are issued together: the raw rows of a u8 sample
    // take 4 registers, of an FP32 sample 16
    constexpr int kSlowBatch = sizeof(TD) == 1 ? (SL < 4 ? SL : 4) : (SL < 2 ? SL : 2);

    // ---- prologue: the subset in registers --------------------------------
    float nref[SL], gxs[SL], gys[SL], val[SL];
    uint64_t refok = 0, inmask = 0, ok = 0;
    const float mean_r = pe->mean, inv_r = pe->inv_std;
    // (q, r) = divmod(tid + T j, dim), advanced by divmod(T, dim) per slot
    const int q64 = T / dim, r64 = T % dim;
    const int q0 = tid / dim, r0 = tid % dim;
    int bad = 0;
    {
        int q = q0, r = r0;
#pragma unroll
        for (int j = 0; j < SL; ++j) {
            nref[j] = 0.0f;
            gxs[j] = 0.0f;
            gys[j] = 0.0f;
            val[j] = 0.0f;
            if (tid + T * j < n) {
                const int a = (t.cy + q - half) * w + t.cx + r - half;
                const float ri = refI[a];
                nref[j] = (ri - mean_r) * inv_r;
                gxs[j] = (gx16 ? half_bits_to_float(gx16[a], bf16) : refGx[a]) * inv_r;
                gys[j] = (gy16 ? half_bits_to_float(gy16[a], bf16) : refGy[a]) * inv_r;
                inmask |= (1ull << j);
                if (!(ri < kGhostWallThreshold)) refok |= (1ull << j);
                else bad++;
            }
            r += r64;
            q += q64;
            if (r >= dim) {
                r -= dim;
                q++;
            }
        }
    }
    const int invalid_ref = group_sum_i<NW>(bad, red, par);
    if (lane < 36) L.H0[lane] = pe->hstep[lane];
    // the FP64 Cholesky step on fully valid iterations, when the pool has one
    const double* ch = chol ? chol + (size_t)t.pool * kCholStride : nullptr;
    const bool use_chol = ch != nullptr && !__builtin_isnan(ch[0]);
    __syncthreads();
    const int min_valid = (int)((float)n * kMinValidPixelFraction);

    // ---- coarse search (estimate_initial_guess, :120-179) ----------------
    float start[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) start[k] = t.g[k];
    if (t.mode == AUTO_SEARCH && t.g[0] == 0.0f && t.g[1] == 0.0f) {
        constexpr int SPL = (7 * 64 + T - 1) / T;   // ceil(ceil(3969 / 9) / T)
        float rr[SPL];
        int sxo[SPL], syo[SPL];
        uint32_t has = 0;
        const int sampled = (n + 8) / 9;
#pragma unroll
        for (int k = 0; k < SPL; ++k) {
            const int i = 9 * (tid + T * k);
            rr[k] = 0.0f;
            sxo[k] = syo[k] = 0;
            if (i < n) {
                sxo[k] = i % dim - half;
                syo[k] = i / dim - half;
                rr[k] = refI[(t.cy + syo[k]) * w + t.cx + sxo[k]];
                if (!(rr[k] < kGhostWallThreshold)) has |= (1u << k);
            }
        }
        float best = 1e20f, bu = 0.0f, bv = 0.0f;
        for (int s = 0; s < 256; ++s) {
            const int u = -15 + 2 * (s & 15), v = -15 + 2 * (s >> 4);
            float dd[SPL];
            uint32_t use = 0;
            float smr = 0.0f, smd = 0.0f;
            int m = 0;
#pragma unroll
            for (int k = 0; k < SPL; ++k) {
                dd[k] = 0.0f;
                if (!((has >> k) & 1)) continue;
                const float x = (float)(t.cx + sxo[k]) + (float)u;
                const float y = (float)(t.cy + syo[k]) + (float)v;
                if (!inside_guard(x, y, w, h)) continue;
                dd[k] = interp<K6>(def, w, x, y);
                use |= (1u << k);
                smr += rr[k];
                smd += dd[k];
                m++;
            }
            // the count with the two sums in one group sum (counts add exactly in float)
            float msd[3] = {smr, smd, (float)m};
            group_sum_n<NW, 3>(msd, red, par);
            m = (int)msd[2];
            if (m < 3 || 2 * m < sampled) continue;
            const float fm = (float)m;
            const float mr = msd[0] / fm, md = msd[1] / fm;
            float vr = 0.0f, vd = 0.0f;
#pragma unroll
            for (int k = 0; k < SPL; ++k)
                if ((use >> k) & 1) {
                    vr += (rr[k] - mr) * (rr[k] - mr);
                    vd += (dd[k] - md) * (dd[k] - md);
                }
            float vv[2] = {vr, vd};
            group_sum_n<NW, 2>(vv, red, par);
            vr = vv[0];
            vd = vv[1];
            const float sr = sqrtf(vr / fm), sdv = sqrtf(vd / fm);
            if (!(sr > 1e-5f) || !(sdv > 1e-5f)) continue;
            float sc = 0.0f;
#pragma unroll
            for (int k = 0; k < SPL; ++k)
                if ((use >> k) & 1) {
                    const float e = (rr[k] - mr) / sr - (dd[k] - md) / sdv;
                    sc += e * e;
                }
            sc = group_sum<NW>(sc, red, par) / fm;
            if (sc < best) {
                best = sc;
                bu = (float)u;
                bv = (float)v;
            }
        }
        start[0] = bu;
        start[1] = bv;
    }

    // ---- the solve as a uniform state machine ------------------------------
    int phase, D = 0, np = 0, s_stage = 0, s_k = 0, s_iter = 0, s_best = 0, s_worst = 0;
    int run = 0;            // 0: first IC-GN run; 1: IC-GN after a rescue
    int simplex = 0;
    float W[9];
    int it = 0;
    float final_score = 1.0f;
    // The Simplex state, the same in every wave, held lane-wise: coordinate
    // j of a point on lane j (j < 6) -- the point under evaluation qv, the
    // centroid, the reflected, the worst and the expanded or contracted
    // point; the vertices in the wave's LDS copy (vertex k, coordinate j at
    // L.P[6 k + j], each lane reading and writing only its own coordinate);
    // and the score of vertex k on lane k (Yl), so that the ranks of all
    // vertices come from one pass of lane-parallel compares.
    float Yl = 0.0f;
    float qv = 0.0f, pbv = 0.0f, prv = 0.0f, pev = 0.0f, pwv = 0.0f;
    float yr = 0.0f, yw = 0.0f, y0 = 0.0f, y2 = 0.0f;
    const int pj = lane < 6 ? lane : 0;   // this lane's coordinate slot
    const auto ldP = [&](int k) -> float { return L.P[k * 6 + pj]; };
    const auto stP = [&](int k, float v) {
        if (lane < 6) L.P[k * 6 + pj] = v;
    };
    const auto lane_f = [&](float v, int j) -> float {
        return __uint_as_float((uint32_t)__builtin_amdgcn_readlane((int)__float_as_uint(v), j));
    };
    const auto setY = [&](int k, float v) { Yl = (lane == k) ? v : Yl; };

    Res first = {}, res = {};
    int trig = 0;

    auto begin_icgn = [&](const float* g) {
        W[0] = 1.0f + g[2];
        W[1] = g[3];
        W[2] = g[0];
        W[3] = g[4];
        W[4] = 1.0f + g[5];
        W[5] = g[1];
        W[6] = 0.0f;
        W[7] = 0.0f;
        W[8] = 1.0f;
        it = 0;
        final_score = 1.0f;
        phase = PH_ICGN;
    };
    auto begin_simplex = [&](const float* s0, int dims) {
        D = dims;
        np = dims + 1;
        float s0v = 0.0f;
#pragma unroll
        for (int j = 0; j < 6; ++j) s0v = (lane == j) ? s0[j] : s0v;
        const float p0 = (lane < D) ? s0v : 0.0f;
#pragma unroll
        for (int k = 1; k < 7; ++k) stP(k, 0.0f);
        stP(0, p0);
        s_stage = S_INIT;
        s_k = 0;
        s_iter = 0;
        qv = p0;
        phase = PH_SIMPLEX;
        simplex = 1;
    };

    if (t.mode == AUTO_SEARCH) {
        begin_simplex(start, 2);
    } else {
        begin_icgn(start);
    }

    // stable ascending order of the vertex scores: rank of each vertex
    auto simplex_iter_begin = [&]() -> bool {
        // rank of vertex `lane` among the np (ties by index), then the vertex of each rank
        int rank = 0;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            const float yk = lane_f(Yl, k);
            rank += (k < np) && ((yk < Yl) || (yk == Yl && k < lane));
        }
        const auto of_rank = [&](int r) -> int {
            return (int)__builtin_ctzll(__builtin_amdgcn_ballot_w64(lane < np && rank == r));
        };
        s_best = of_rank(0);
        s_worst = of_rank(np - 1);
        y0 = lane_f(Yl, s_best);
        yw = lane_f(Yl, s_worst);
        y2 = lane_f(Yl, of_rank(np - 2));
        if (fabsf(y0 - yw) < 1e-5f) return false;
        // the centroid of all but the worst, summed in rank order
        const bool in_dim = lane < D;
        pbv = 0.0f;
        for (int r = 0; r < D; ++r) {
            const float v = ldP(of_rank(r));
            if (in_dim) pbv += v;
        }
        if (in_dim) pbv /= (float)D;
        pwv = ldP(s_worst);
        prv = in_dim ? pbv + 1.0f * (pbv - pwv) : 0.0f;
        qv = prv;
        s_stage = S_REFLECT;
        return true;
    };
    auto replace_worst = [&](float ptv, float yv) {
        if (lane < D) stP(s_worst, ptv);
        setY(s_worst, yv);
    };
    // Simplex done: best vertex, then the IC-GN run it seeds.
    auto finish_simplex = [&]() {
        int best = 0;
        float yb = lane_f(Yl, 0);
        for (int k = 1; k < np; ++k) {
            const float yk = lane_f(Yl, k);
            if (yk < yb) {
                best = k;
                yb = yk;
            }
        }
        const float bv = ldP(best);
        float p[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) p[j] = (j < D) ? lane_f(bv, j) : 0.0f;
        begin_icgn(p);
    };
    // one IC-GN run ended with r
    auto finish_icgn = [&](const Res& r) {
        if (run == 0) {
            res = r;
            trig = r.iters;
            const bool rescue = t.mode == NO_SEARCH && (r.status != 0 || r.corr > kCorrSimplexTrigger);
            if (rescue) {
                first = r;
                run = 1;
                begin_simplex(t.g, 6);
            } else {
                phase = PH_DONE;
            }
        } else {
            const bool first_better = first.status == 0 && (r.status != 0 || first.corr < r.corr);
            res = first_better ? first : r;
            phase = PH_DONE;
        }
    };
    auto icgn_result = [&](int status, float score, int iters) {
        Res r;
        r.p[0] = W[2];
        r.p[1] = W[5];
        r.p[2] = W[0] - 1.0f;
        r.p[3] = W[1];
        r.p[4] = W[3];
        r.p[5] = W[4] - 1.0f;
        r.corr = score;
        r.status = status;
        r.iters = iters;
        return r;
    };

    while (phase != PH_DONE) {
        const bool icgn = phase == PH_ICGN;
        float Wc[6];
        if (icgn) {
#pragma unroll
            for (int k = 0; k < 6; ++k) Wc[k] = W[k];
        } else {
            float q[6];
#pragma unroll
            for (int j = 0; j < 6; ++j) q[j] = lane_f(qv, j);
            Wc[0] = 1.0f + q[2];
            Wc[1] = q[3];
            Wc[2] = q[0];
            Wc[3] = q[4];
            Wc[4] = 1.0f + q[5];
            Wc[5] = q[1];
        }

        // ---- the single sampling site -------------------------------------
        float sum = 0.0f;
        int cnt = 0;
        ok = 0;
        {
            // The slot coordinates are recomputed every step: left loop-invariant,
            // the compiler hoists per-pixel products and doubles out of the solve
            // loop and the register file, not the ALUs, becomes the limit.
            int qq = q0, rq = r0;
            asm volatile("" : "+v"(qq), "+v"(rq));
            if constexpr (!K6) {
                // Keys 4x4 in two phases (k4_issue, then the reference's
                // arithmetic on the loaded rows), SB slots' gathers in flight
                // at once; a slot off the subset or outside the guard loads
                // the subset's centre and keeps a zero.
#pragma unroll
                for (int j0 = 0; j0 < SL; j0 += kSlowBatch) {
                    K4Rows kr[kSlowBatch];
                    bool inb[kSlowBatch];
#pragma unroll
                    for (int b = 0; b < kSlowBatch; ++b) {
                        const int j = j0 + b;
                        inb[b] = false;
                        if (j < SL) {
                            const float x = (float)(rq - half), y = (float)(qq - half);
                            const float fx = warped(t.cx, Wc[0], Wc[1], Wc[2], x, y);
                            const float fy = warped(t.cy, Wc[3], Wc[4], Wc[5], x, y);
                            inb[b] = ((inmask >> j) & 1) && inside_guard(fx, fy, w, h);
                            k4_issue<TD, false>(drs, w, inb[b] ? fx : (float)t.cx, inb[b] ? fy : (float)t.cy, kr[b]);
                            rq += r64;
                            qq += q64;
                            if (rq >= dim) {
                                rq -= dim;
                                qq++;
                            }
                        }
                    }
#pragma unroll
                    for (int b = 0; b < kSlowBatch; ++b) {
                        const int j = j0 + b;
                        if (j >= SL) continue;
                        val[j] = 0.0f;
                        if (inb[b]) {
                            const float v = k4_eval_ref<TD>(kr[b]);
                            val[j] = v;
                            if ((refok >> j) & 1) {
                                ok |= (1ull << j);
                                sum += v;
                                cnt++;
                            }
                        }
                    }
                }
            } else {
#pragma unroll
                for (int j = 0; j < SL; ++j) {
                    val[j] = 0.0f;
                    if ((inmask >> j) & 1) {
                        const float x = (float)(rq - half), y = (float)(qq - half);
                        const float fx = warped(t.cx, Wc[0], Wc[1], Wc[2], x, y);
                        const float fy = warped(t.cy, Wc[3], Wc[4], Wc[5], x, y);
                        if (inside_guard(fx, fy, w, h)) {
                            const float v = interp<K6>(def, w, fx, fy);
                            val[j] = v;
                            if ((refok >> j) & 1) {
                                ok |= (1ull << j);
                                sum += v;
                                cnt++;
                            }
                        }
                    }
                    rq += r64;
                    qq += q64;
                    if (rq >= dim) {
                        rq -= dim;
                        qq++;
                    }
                }
            }
        }
        // the sum and the count in one group sum (counts add exactly in float)
        float sc[2] = {sum, (float)cnt};
        group_sum_n<NW, 2>(sc, red, par);
        sum = sc[0];
        const int nv = (int)sc[1];

        if (!icgn) {
            // ---- evaluate_znssd (:467-528) ---------------------------------
            float score = 2.0f;
            if (nv != 0 && nv >= min_valid) {
                const float mean = sum / (float)nv;
                float s2 = 0.0f;
#pragma unroll
                for (int j = 0; j < SL; ++j)
                    if ((ok >> j) & 1) {
                        const float d = val[j] - mean;
                        s2 += d * d;
                    }
                float dstd = sqrtf(group_sum<NW>(s2, red, par) / (float)nv);
                if (dstd < 1e-5f) dstd = 1.0f;
                const float inv = 1.0f / dstd;
                const bool full = nv == n;
                float z = 0.0f;
#pragma unroll
                for (int j = 0; j < SL; ++j)
                    if ((ok >> j) & 1) {
                        const float nd = full ? (val[j] - mean) * inv : (val[j] - mean) / dstd;
                        const float d = nref[j] - nd;
                        z += d * d;
                    }
                score = group_sum<NW>(z, red, par) / (float)nv;
                if (!finite_f(score)) score = 2.0f;
            }
            // ---- Nelder-Mead step (:531-651) --------------------------------
            bool next_iter = false;
            const bool in_dim = lane < D;
            if (s_stage == S_INIT) {
                setY(s_k, score);
                s_k++;
                if (s_k < np) {
                    // vertex s_k: vertex 0 with coordinate s_k - 1 moved by its scale
                    float v = ldP(0);
                    if (lane == s_k - 1) v += (lane < 2) ? 2.0f : 0.01f;
                    const float pk = in_dim ? v : 0.0f;
                    stP(s_k, pk);
                    qv = pk;
                } else {
                    if (!simplex_iter_begin()) finish_simplex();
                }
            } else if (s_stage == S_REFLECT) {
                yr = score;
                if (y0 <= yr && yr < y2) {
                    replace_worst(prv, yr);
                    next_iter = true;
                } else if (yr < y0) {
                    pev = in_dim ? pbv + 2.0f * (prv - pbv) : 0.0f;
                    qv = pev;
                    s_stage = S_EXPAND;
                } else {
                    const bool outside = yr < yw;
                    pev = in_dim ? pbv + 0.5f * ((outside ? prv : pwv) - pbv) : 0.0f;
                    qv = pev;
                    s_stage = S_CONTRACT;
                }
            } else if (s_stage == S_EXPAND) {
                if (score < yr) replace_worst(pev, score);
                else replace_worst(prv, yr);
                next_iter = true;
            } else if (s_stage == S_CONTRACT) {
                if (score < fminf(yr, yw)) {
                    replace_worst(pev, score);
                    next_iter = true;
                } else {
                    // shrink every vertex but the best towards it, one evaluation each
                    s_k = (s_best == 0) ? 1 : 0;
                    const float b = ldP(s_best);
                    float pk = ldP(s_k);
                    if (in_dim) pk = b + 0.5f * (pk - b);
                    stP(s_k, pk);
                    qv = pk;
                    s_stage = S_SHRINK;
                }
            } else {  // S_SHRINK
                setY(s_k, score);
                int nk = s_k + 1;
                if (nk == s_best) nk++;
                if (nk < np) {
                    s_k = nk;
                    const float b = ldP(s_best);
                    float pk = ldP(s_k);
                    if (in_dim) pk = b + 0.5f * (pk - b);
                    stP(s_k, pk);
                    qv = pk;
                } else {
                    next_iter = true;
                }
            }
            if (next_iter) {
                s_iter++;
                if (s_iter >= 100 || !simplex_iter_begin()) finish_simplex();
            }
            continue;
        }

        // ---- IC-GN iteration (solve_icgn, :332-462) ---------------------------
        if (nv < min_valid) {
            finish_icgn(icgn_result(1, 2.0f, it));
            continue;
        }
        const float mean = sum / (float)nv;
        float s2 = 0.0f;
#pragma unroll
        for (int j = 0; j < SL; ++j)
            if ((ok >> j) & 1) {
                const float d = val[j] - mean;
                s2 += d * d;
            }
        float dstd = sqrtf(group_sum<NW>(s2, red, par) / (float)nv);
        if (dstd < 1e-5f) dstd = 1.0f;
        const float inv = 1.0f / dstd;
        float err = 0.0f, dp[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        const bool masked = nv != n;
        float hh[21];
#pragma unroll
        for (int k = 0; k < 21; ++k) hh[k] = 0.0f;
        {
            int qq = q0, rq = r0;
            asm volatile("" : "+v"(qq), "+v"(rq));
#pragma unroll
            for (int j = 0; j < SL; ++j) {
                if ((ok >> j) & 1) {
                    const float x = (float)(rq - half), y = (float)(qq - half);
                    const float nd = (val[j] - mean) * inv;
                    const float d = nref[j] - nd;
                    const float s6[6] = {gxs[j], gys[j], gxs[j] * x, gxs[j] * y, gys[j] * x, gys[j] * y};
                    err += d * d;
#pragma unroll
                    for (int k = 0; k < 6; ++k) dp[k] += s6[k] * d;
                    if (masked) {
                        int k = 0;
#pragma unroll
                        for (int a = 0; a < 6; ++a)
#pragma unroll
                            for (int b = a; b < 6; ++b) hh[k++] += s6[a] * s6[b];
                    }
                }
                rq += r64;
                qq += q64;
                if (rq >= dim) {
                    rq -= dim;
                    qq++;
                }
            }
        }
        {
            float ed[7] = {err, dp[0], dp[1], dp[2], dp[3], dp[4], dp[5]};
            group_sum_n<NW, 7>(ed, red, par);
            err = ed[0];
#pragma unroll
            for (int k = 0; k < 6; ++k) dp[k] = ed[1 + k];
        }
        float d[6];
        if (masked) {
            group_sum_n<NW, 21>(hh, red, par);
            // masked Hessian (:373-413): 2x2 test, LM, rcond, inverse, with the
            // 6x6 spread one entry per lane (lane 6 r + c) and the same
            // operations per entry as the pool precompute's register routines
            const int hr = lane / 6, hc = lane - 6 * (lane / 6);
            float hL;
            {
                // entry (r, c) of the symmetric matrix is sum k of its upper triangle
                const int a0 = hr < hc ? hr : hc, b0 = hr < hc ? hc : hr;
                const int tk = a0 * 6 - (a0 * (a0 - 1)) / 2 + (b0 - a0);
                hL = hh[0];
#pragma unroll
                for (int k = 1; k < 21; ++k) hL = (tk == k) ? hh[k] : hL;
            }
            const float det2 = hh[0] * hh[6] - hh[1] * hh[1];
            const float nrm2 = hh[0] * hh[0] + hh[1] * hh[1] + hh[1] * hh[1] + hh[6] * hh[6];
            const float cond2 = (fabsf(det2) > 1e-12f) ? (nrm2 / fabsf(det2)) : 1.0e13f;
            bool flag = cond2 > 1.0e12f;
            float hmv = 0.0f;   // row `lane` of the step matrix times dp, on lanes 0..5
            if (!flag) {
                if (lm && (lane == 0 || lane == 7)) hL += kLmAlpha;
                if constexpr (NW > 1) {
                    // the rcond test on wave 0 and the inverse on wave 1 at once
                    // (the same sequences, so the same bits); every wave takes
                    // both from the exchange slots after the group-sum halves,
                    // which the next write reaches only past later group sums
                    float* xch = red + 2 * NW * kGroupSlots;
                    const int wid = (int)(threadIdx.x >> 6);
                    if (wid == 0) {
                        const bool wc = well_conditioned_lanes(hL, half, lane);
                        if (lane == 0) xch[36] = wc ? 1.0f : 0.0f;
                    } else if (wid == 1) {
                        float mL = hL, mR = (hr == hc) ? 1.0f : 0.0f;
                        gj_inverse_lanes(mL, mR, lane);
                        if (lane < 36) xch[lane] = mR;
                    }
                    __syncthreads();
                    flag = xch[36] == 0.0f;
                    if (!flag) {
                        const int rr = lane < 6 ? lane : 0;
                        float acc = 0.0f;
#pragma unroll
                        for (int c = 0; c < 6; ++c) acc += xch[rr * 6 + c] * dp[c];
                        hmv = -acc;
                    }
                } else {
                    flag = !well_conditioned_lanes(hL, half, lane);
                    if (!flag) {
                        float mL = hL, mR = (hr == hc) ? 1.0f : 0.0f;
                        gj_inverse_lanes(mL, mR, lane);
                        const int rr = lane < 6 ? lane : 0;
                        float acc = 0.0f;
#pragma unroll
                        for (int c = 0; c < 6; ++c) acc += lane_fetch(mR, rr * 6 + c) * dp[c];
                        hmv = -acc;
                    }
                }
            }
            if (flag) {
                finish_icgn(icgn_result(1, 2.0f, it));
                continue;
            }
            final_score = err / (float)nv;
#pragma unroll
            for (int r = 0; r < 6; ++r) d[r] = lane_read(hmv, r);
        } else {
            final_score = err / (float)nv;
            if (use_chol) {
                chol64_step(ch, dp, d);
            } else {
#pragma unroll
                for (int r = 0; r < 6; ++r) {
                    float acc = 0.0f;
#pragma unroll
                    for (int c = 0; c < 6; ++c) acc += L.H0[r * 6 + c] * dp[c];
                    d[r] = -acc;
                }
            }
        }
        const float dW[9] = {1.0f + d[2], d[3], d[0], d[4], 1.0f + d[5], d[1], 0.0f, 0.0f, 1.0f};
        const float det = dW[0] * dW[4] - dW[1] * dW[3];
        bool fin = true;
#pragma unroll
        for (int k = 0; k < 6; ++k) fin = fin && finite_f(d[k]);
        if (!fin || !(fabsf(det) > 1e-6f)) {
            finish_icgn(icgn_result(1, 2.0f, it + 1));
            continue;
        }
        float dWi[9], Wn[9];
        inv3(dW, dWi);
        mul3(W, dWi, Wn);
#pragma unroll
        for (int k = 0; k < 9; ++k) fin = fin && finite_f(Wn[k]);
        if (!fin) {
            finish_icgn(icgn_result(1, 2.0f, it + 1));
            continue;
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) W[k] = Wn[k];
        float nrm = 0.0f;
#pragma unroll
        for (int k = 0; k < 6; ++k) nrm += d[k] * d[k];
        if (sqrtf(nrm) < conv_tol) {
            // texture check over the surviving pixels (:436-459)
            float sx = 0.0f, sy = 0.0f;
#pragma unroll
            for (int j = 0; j < SL; ++j)
                if ((ok >> j) & 1) {
                    float a = gxs[j], b = gys[j];
                    asm volatile("" : "+v"(a), "+v"(b));
                    sx += a * a;
                    sy += b * b;
                }
            float sxy[2] = {sx, sy};
            group_sum_n<NW, 2>(sxy, red, par);
            sx = sxy[0];
            sy = sxy[1];
            if (fminf(sx, sy) <= 0.0f || !finite_f(final_score))
                finish_icgn(icgn_result(1, 2.0f, it + 1));
            else
                finish_icgn(icgn_result(0, final_score, it + 1));
            continue;
        }
        it++;
        if (it >= kIcgnMaxIter) finish_icgn(icgn_result(1, final_score, kIcgnMaxIter));
    }

    if (lm && (float)invalid_ref / (float)(dim * dim) > kGhostRejectFraction) {
        res.corr = 2.0f;
        res.status = 1;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) out.p[k] = res.p[k];
    out.corr = res.corr;
    out.status = res.status;
    out.iters = res.iters;
    out.trigger_iters = trig;
    out.simplex = simplex;
    out.init = 1;
    out.invalid_ref = invalid_ref;
}


// The common case, lean: IC-GN from the guess on a subset whose every pixel
// samples inside the image and off the Ghost Wall, with the pool's step matrix
// (no masked Hessian, no Simplex, no coarse search). Returns false -- the work
// item goes to the full solver, which starts it again from scratch -- as soon
// as a masked iteration, a Simplex rescue or a coarse search would be needed,
// so the committed results are exactly full_solve's.
template <int PPL, bool K6, typename TD, bool COEF>
__device__ __forceinline__ bool fast_solve(const Task& t, const PoolEntry* __restrict__ pool,
                                           const float* __restrict__ refI,
                                           const float* __restrict__ refGx,
                                           const float* __restrict__ refGy,
                                           const uint16_t* __restrict__ gx16, const uint16_t* __restrict__ gy16, bool bf16,
                                           const double* __restrict__ chol,
                                           const TD* __restrict__ def, const uint4* __restrict__ defc,
                                           int w, int h, int dim, bool lm, float stop2, SolveOut& out) {
    const int lane = threadIdx.x & 63;
    const PoolEntry* pe = pool + __builtin_amdgcn_readfirstlane(t.pool);
    out.init = 0;
    out.simplex = 0;
    if (!pe->valid) return true;
    const int half = dim / 2;
    if (t.cx - half < 0 || t.cx + half >= w || t.cy - half < 0 || t.cy + half >= h) return true;
    if (t.mode == AUTO_SEARCH) return false;
    const int n = dim * dim;
    float nref[PPL], gxs[PPL], gys[PPL], val[PPL];
    sdic_h2 xy[PPL];   // the pixel's offset from the centre, (x, y), exact in fp16
    const float mean_r = pe->mean, inv_r = pe->inv_std;
    const int q64 = 64 / dim, r64 = 64 % dim;
    // pixels j < kFull exist in every lane: launch_solve picks the smallest
    // class, so the subset has more pixels than 64 lanes of the next smaller
    constexpr int kFull = PPL == 4 ? 0 : PPL == 8 ? 4 : PPL == 12 ? 8 : PPL == 16 ? 12 : PPL == 24 ? 16
                        : PPL == 32 ? 24 : PPL == 48 ? 32 : 48;
    int bad = 0;
    {
        // Branch-free, so that all 3 PPL loads are in flight together: a lane
        // past the subset's last pixel reads the centre (and keeps zeros) and
        // samples the centre in the iterations (its sample is left out of
        // every sum).
        // Pixel (q, r) of the subset (row q, column r) is q w + r past its
        // top-left pixel; the offsets advance by additions only.
        int at[PPL];
        int q = lane / dim, r = lane - (lane / dim) * dim;
        const int ctr = t.cy * w + t.cx;
        const int adv = q64 * w + r64, wrap = w - dim;
        int off = ctr - half * w - half + q * w + r;
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
            const bool in = j < kFull || lane + 64 * j < n;
            xy[j] = in ? sdic_h2{(_Float16)(r - half), (_Float16)(q - half)} : sdic_h2{0, 0};
            at[j] = in ? off : ctr;
            r += r64;
            q += q64;
            off += adv;
            if (r >= dim) {
                r -= dim;
                q++;
                off += wrap;
            }
        }
        // the three planes through buffer resources: one 32-bit byte offset per
        // pixel serves all three loads
        const int plane_bytes = (int)((size_t)w * h * sizeof(float));
        const __amdgpu_buffer_rsrc_t rI = __builtin_amdgcn_make_buffer_rsrc((void*)refI, (short)0, plane_bytes, 0x00020000);
#pragma unroll
        for (int j = 0; j < PPL; ++j) nref[j] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rI, at[j] * 4, 0, 0));
        if (gx16) {
#pragma unroll
            for (int j = 0; j < PPL; ++j) {
                gxs[j] = half_bits_to_float(gx16[at[j]], bf16);
                gys[j] = half_bits_to_float(gy16[at[j]], bf16);
            }
        } else {
            const __amdgpu_buffer_rsrc_t rX =
                __builtin_amdgcn_make_buffer_rsrc((void*)refGx, (short)0, plane_bytes, 0x00020000);
            const __amdgpu_buffer_rsrc_t rY =
                __builtin_amdgcn_make_buffer_rsrc((void*)refGy, (short)0, plane_bytes, 0x00020000);
#pragma unroll
            for (int j = 0; j < PPL; ++j) {
                gxs[j] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rX, at[j] * 4, 0, 0));
                gys[j] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rY, at[j] * 4, 0, 0));
            }
        }
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
            const bool in = j < kFull || lane + 64 * j < n;
            const float ri = nref[j];
            bad += (in && ri < kGhostWallThreshold) ? 1 : 0;
            nref[j] = in ? (ri - mean_r) * inv_r : 0.0f;
            gxs[j] = in ? gxs[j] * inv_r : 0.0f;
            gys[j] = in ? gys[j] * inv_r : 0.0f;
            val[j] = 0.0f;
        }
    }
    // Any Ghost Wall pixel makes every iteration a masked one.
    if (wave_sum_i(bad) != 0) return false;
    const float* __restrict__ hs = pe->hstep;   // wave-uniform: scalar loads
    // Lane-parallel pieces of the per-iteration bookkeeping: lane l holds row
    // (l & 7) of the step matrix (rows 6 and 7 repeat row 0) and tests warped
    // corner (l & 3) of the subset.
    float hrow[6];
    {
        const int hr = (lane & 7) < 6 ? (lane & 7) : 0;
#pragma unroll
        for (int c = 0; c < 6; ++c) hrow[c] = hs[hr * 6 + c];
    }
    const double* __restrict__ ch =
        chol ? chol + (size_t)__builtin_amdgcn_readfirstlane(t.pool) * kCholStride : nullptr;
    const bool use_chol = ch != nullptr && !__builtin_isnan(ch[0]);
    // the texture test at convergence: min(sum gx^2, sum gy^2) > 0 over the subset
    float tgx = 0.0f, tgy = 0.0f;
#pragma unroll
    for (int j = 0; j < PPL; ++j) {
        tgx = __builtin_fmaf(gxs[j], gxs[j], tgx);
        tgy = __builtin_fmaf(gys[j], gys[j], tgy);
    }
    const bool textured = fminf(wave_sum(tgx), wave_sum(tgy)) > 0.0f;
    const int min_valid = (int)((float)n * kMinValidPixelFraction);
    const double cxd = (double)t.cx, cyd = (double)t.cy, hd = (double)half;
    const float cxf = (float)t.cx, cyf = (float)t.cy, hf = (float)half;
    // a / n correctly rounded without a division (a sum over the subset's n
    // pixels; a sampled iteration reaching the sums has all n valid): with
    // rn = RN(1 / n), q = RN(a rn) is within an ulp of a / n, the residual
    // a - q n is exact as an FMA, and q + (a - q n) rn rounds to RN(a / n),
    // the same bits as the division (checked exhaustively for every odd side
    // 3 .. 83 and every non-negative finite float a; the sums are never negative)
    const float fn = (float)n, rn = 1.0f / fn;
    const auto div_n = [&](float a) {
        const float q = a * rn;
        return __builtin_fmaf(__builtin_fmaf(-q, fn, a), rn, q);
    };
    const __amdgpu_buffer_rsrc_t rs =
        COEF ? __builtin_amdgcn_make_buffer_rsrc((void*)defc, (short)0, (int)((size_t)w * h * 16), 0x00020000)
             : __builtin_amdgcn_make_buffer_rsrc((void*)def, (short)0, (int)((size_t)w * h * sizeof(TD)), 0x00020000);
    float W[9] = {1.0f + t.g[2], t.g[3], t.g[0], t.g[4], 1.0f + t.g[5], t.g[1], 0.0f, 0.0f, 1.0f};
    float final_score = 1.0f;
    int status = 1, iters = kIcgnMaxIter;
    float corr = 0.0f;
    bool done = false;
    for (int it = 0; it < kIcgnMaxIter && !done; ++it) {
        const double w0 = W[0], w1 = W[1], w2 = W[2], w3 = W[3], w4 = W[4], w5 = W[5];
        // Interior: the four warped corners lie 1 px inside the 4 px guard, so
        // (the map being affine) every pixel does, and no pixel test is needed.
        // each lane tests one of the four corners; interior when none is out.
        // The corners in FP32 first: where they are within the frame they are
        // within 0.02 px of the FP64 corners, so corners 1 px further inside
        // than the test asks pass it in FP64 too, and only a subset near the
        // guard takes the FP64 test.
        bool interior;
        {
            const float xs = (lane & 1) ? hf : -hf, ys = (lane & 2) ? hf : -hf;
            const float gx = cxf + __builtin_fmaf(W[0], xs, __builtin_fmaf(W[1], ys, W[2]));
            const float gy = cyf + __builtin_fmaf(W[3], xs, __builtin_fmaf(W[4], ys, W[5]));
            const bool sure = gx >= 5.5f && gx < (float)(w - 7) && gy >= 5.5f && gy < (float)(h - 7);
            interior = __builtin_amdgcn_ballot_w64(!sure) == 0;
        }
        if (!interior) {
            const double x = (lane & 1) ? hd : -hd, y = (lane & 2) ? hd : -hd;
            const float fx = warped_fma(cxd, w0, w1, w2, x, y), fy = warped_fma(cyd, w3, w4, w5, x, y);
            const bool in = fx + 0.5f >= 5.0f && fx + 0.5f < (float)(w - 5) && fy + 0.5f >= 5.0f &&
                            fy + 0.5f < (float)(h - 5);
            interior = __builtin_amdgcn_ballot_w64(!in) == 0;
        }
        float sum = 0.0f;
        int nv = n;
        if (interior) {
            // Inside the margin the warp is evaluated in FP32: it can differ from
            // the FP64 position by one float ulp, far inside the 1 px margin.
            // No branch per pixel, and (Keys 4x4) the loads of kBatch pixels
            // issued before any is evaluated, so their gathers overlap.
            if constexpr (!K6) {
                constexpr int kBatch = (COEF || sizeof(TD) == 1) ? SDIC_K4_BATCH_NARROW : SDIC_K4_BATCH_F32;
                static_assert(PPL % kBatch == 0, "batch must divide the pixels per lane");
#pragma unroll
                for (int j0 = 0; j0 < PPL; j0 += kBatch) {
                    K4Rows kr[kBatch];
#pragma unroll
                    for (int b = 0; b < kBatch; ++b) {
                        sdic_h2 c = xy[j0 + b];
                        asm volatile("" : "+v"(c));
                        const float x = (float)c.x, y = (float)c.y;
                        const float fx = cxf + __builtin_fmaf(W[0], x, __builtin_fmaf(W[1], y, W[2]));
                        const float fy = cyf + __builtin_fmaf(W[3], x, __builtin_fmaf(W[4], y, W[5]));
                        k4_issue<TD, COEF>(rs, w, fx, fy, kr[b]);
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int b = 0; b < kBatch; ++b) {
                        const int j = j0 + b;
                        const float v = k4_eval<TD, COEF>(kr[b]);
                        val[j] = v;
                        sum += (j < kFull || lane + 64 * j < n) ? v : 0.0f;
                    }
                }
            } else {
#pragma unroll
                for (int j = 0; j < PPL; ++j) {
                    sdic_h2 c = xy[j];
                    asm volatile("" : "+v"(c));
                    const float x = (float)c.x, y = (float)c.y;
                    const float fx = cxf + __builtin_fmaf(W[0], x, __builtin_fmaf(W[1], y, W[2]));
                    const float fy = cyf + __builtin_fmaf(W[3], x, __builtin_fmaf(W[4], y, W[5]));
                    const float v = interp_buf<K6, TD>(rs, w, fx, fy);
                    val[j] = v;
                    sum += lane + 64 * j < n ? v : 0.0f;
                }
            }
        } else {
            int cnt = 0;
#pragma unroll
            for (int j = 0; j < PPL; ++j) {
                val[j] = 0.0f;
                if (lane + 64 * j < n) {
                    sdic_h2 c = xy[j];
                    asm volatile("" : "+v"(c));
                    const double x = (double)(float)c.x, y = (double)(float)c.y;
                    const float fx = warped_fma(cxd, w0, w1, w2, x, y);
                    const float fy = warped_fma(cyd, w3, w4, w5, x, y);
                    if (inside_guard(fx, fy, w, h)) {
                        const float v = interp_fma<K6>(def, w, fx, fy);
                        val[j] = v;
                        sum += v;
                        cnt++;
                    }
                }
            }
            nv = wave_sum_i(cnt);
            if (nv < min_valid) {
                status = 1;
                corr = 2.0f;
                iters = it;
                done = true;
                break;
            }
            if (nv != n) return false;   // a masked iteration
        }
        sum = wave_sum(sum);
        const float mean = div_n(sum);
        float s2 = 0.0f;
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
            const float d = val[j] - mean;
            s2 = (j < kFull || lane + 64 * j < n) ? __builtin_fmaf(d, d, s2) : s2;
        }
        float dstd = sqrtf(div_n(wave_sum(s2)));
        if (dstd < 1e-5f) dstd = 1.0f;
        const float inv = 1.0f / dstd;
        float dp[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        // The ZNSSD of this iteration's sample, sum d^2 / n: only the iteration
        // that reports it (the converged or the last one) forms it, in the
        // same order as the residual loop below.
        const auto znssd = [&]() {
            float e = 0.0f;
#pragma unroll
            for (int j = 0; j < PPL; ++j) {
                const float d = (j < kFull || lane + 64 * j < n) ? nref[j] - (val[j] - mean) * inv : 0.0f;
                e = __builtin_fmaf(d, d, e);
            }
            return div_n(wave_sum(e));
        };
        // branch-free: a lane past the last pixel has d = 0 (and zero
        // gradients and offsets), so it adds exact zeros
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
            sdic_h2 c = xy[j];
            asm volatile("" : "+v"(c));
            const float x = (float)c.x, y = (float)c.y;
            const float d = (j < kFull || lane + 64 * j < n) ? nref[j] - (val[j] - mean) * inv : 0.0f;
            const float gd = gxs[j] * d, hdd = gys[j] * d;
            dp[0] += gd;
            dp[1] += hdd;
            dp[2] = __builtin_fmaf(gd, x, dp[2]);
            dp[3] = __builtin_fmaf(gd, y, dp[3]);
            dp[4] = __builtin_fmaf(hdd, x, dp[4]);
            dp[5] = __builtin_fmaf(hdd, y, dp[5]);
        }
        wave_sum6(dp);
        float d[6];
        if (use_chol) {
            chol64_step(ch, dp, d);
        } else {
            // row (lane & 7) of the step on each lane, in the order of the
            // row-by-row product, then the six rows read back
            float acc = 0.0f;
#pragma unroll
            for (int c = 0; c < 6; ++c) acc += hrow[c] * dp[c];
            const float dr = -acc;
#pragma unroll
            for (int r = 0; r < 6; ++r)
                d[r] = __uint_as_float((uint32_t)__builtin_amdgcn_readlane((int)__float_as_uint(dr), r));
        }
        const float dW[9] = {1.0f + d[2], d[3], d[0], d[4], 1.0f + d[5], d[1], 0.0f, 0.0f, 1.0f};
        const float det = dW[0] * dW[4] - dW[1] * dW[3];
        bool fin = true;
#pragma unroll
        for (int k = 0; k < 6; ++k) fin = fin && finite_f(d[k]);
        if (!fin || !(fabsf(det) > 1e-6f)) {
            status = 1;
            corr = 2.0f;
            iters = it + 1;
            done = true;
            break;
        }
        float Wn[9];
        compose_inverse_affine(W, dW, Wn);
#pragma unroll
        for (int k = 0; k < 9; ++k) fin = fin && finite_f(Wn[k]);
        if (!fin) {
            status = 1;
            corr = 2.0f;
            iters = it + 1;
            done = true;
            break;
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) W[k] = Wn[k];
        float nrm = 0.0f;
#pragma unroll
        for (int k = 0; k < 6; ++k) nrm += d[k] * d[k];
        if (nrm < stop2) {   // sqrtf(nrm) < tol
            iters = it + 1;
            final_score = znssd();
            if (!textured || !finite_f(final_score)) {
                status = 1;
                corr = 2.0f;
            } else {
                status = 0;
                corr = final_score;
            }
            done = true;
            break;
        }
        if (it == kIcgnMaxIter - 1) final_score = znssd();
    }
    if (!done) {
        status = 1;
        corr = final_score;
        iters = kIcgnMaxIter;
    }
    // the rescue this mode allows is the full solver's business
    if (t.mode == NO_SEARCH && (status != 0 || corr > kCorrSimplexTrigger)) return false;
    out.p[0] = W[2];
    out.p[1] = W[5];
    out.p[2] = W[0] - 1.0f;
    out.p[3] = W[1];
    out.p[4] = W[3];
    out.p[5] = W[4] - 1.0f;
    out.corr = corr;
    out.status = status;
    out.iters = iters;
    out.trigger_iters = iters;
    out.simplex = 0;
    out.init = 1;
    out.invalid_ref = 0;
    return true;
}

}  // namespace

// Work item k of a launch: a task (task mode, a.tasks != null) or a grid node
// (Path A, full_field_path_a.cpp:28-114), through the optional index list.
__device__ __forceinline__ bool item_task(const SolveArgs& a, int id, Task& t) {
    if (a.tasks) {
        t = a.tasks[id];
        return true;
    }
    const GridArgs& g = a.grid;
    // the three bytes read together (one memory round trip, not three)
    const uint8_t in_mesh = g.in_mesh[id], solved = g.grid.solved[id];
    const int tag = g.grid.tag[id];
    if (!in_mesh || solved) return false;
    if (g.skip_model && tag == kMeshHomography) return false;
    t.cx = g.rect_x + (id % g.gw) * g.step;
    t.cy = g.rect_y + (id / g.gw) * g.step;
    t.pool = id;
    t.mode = (tag == kMeshExtrapolated || tag == kMeshHomography || g.no_rescue) ? NO_SIMPLEX : NO_SEARCH;
    t.g[0] = g.gu[id];
    t.g[1] = g.gv[id];
    t.g[2] = g.gux[id];
    t.g[3] = g.guy[id];
    t.g[4] = g.gvx[id];
    t.g[5] = g.gvy[id];
    t.node = id;
    return true;
}

// Lane 0 commits a work item: the result slot in task mode; in grid mode the
// accepted node into the grid, and the Path A counters (points, iterations,
// Simplex calls / saved / dead) with integer atomics into the workgroup's
// counter slot.
__device__ __forceinline__ void commit(const SolveArgs& a, int id, const SolveOut& o) {
    if (a.tasks) {
        a.out[id] = o;
        return;
    }
    if (!o.init) return;
    GridDev gd = a.grid.grid;
    const bool acc = o.status == 0 && o.corr <= kCorrAccept;
    int* const st = a.stats + (blockIdx.x % kStatSlots) * kStatStride;
    atomicAdd(&st[1], o.iters);
    if (o.simplex) {
        atomicAdd(&st[2], 1);
        atomicAdd(acc ? &st[3] : &st[4], 1);
    }
    gd.iters[id] = o.iters;
    gd.simplex[id] = o.simplex ? 1 : 0;
    if (acc) {
        gd.u[id] = o.p[0];
        gd.v[id] = o.p[1];
        gd.ux[id] = o.p[2];
        gd.uy[id] = o.p[3];
        gd.vx[id] = o.p[4];
        gd.vy[id] = o.p[5];
        gd.corr[id] = o.corr;
        gd.solved[id] = 1;
        atomicAdd(&st[0], 1);
    } else {
        gd.corr[id] = kCorrInvalid;
    }
}

// Grid order of Path A's work items in tiles of T x T nodes (row-major tiles,
// row-major within a tile; the last tile row and column narrower): a
// bijection of 0 .. gw gh - 1, so that the consecutive items a workgroup band
// takes on one XCD are squares of nodes sharing their subsets' rows.
__device__ __forceinline__ int tile_order(int k, int gw, int gh, int T) {
    const int band = T * gw;
    const int ty = k / band, r = k - ty * band;
    const int th = min(T, gh - ty * T);
    const int full = (gw / T) * th * T;
    int tx, tw, local;
    if (r < full) {
        tx = r / (th * T);
        tw = T;
        local = r - tx * th * T;
    } else {
        tx = gw / T;
        tw = gw - tx * T;
        local = r - full;
    }
    const int iy = local / tw, ix = local - iy * tw;
    return (ty * T + iy) * gw + tx * T + ix;
}

// Fast path, one wave per work item (a.wpb items per workgroup); items it
// cannot finish are appended to a.bail for k_slow.
#ifdef SDIC_FAST_MIN_WAVES
#define SDIC_FAST_ATTR __attribute__((amdgpu_waves_per_eu(SDIC_FAST_MIN_WAVES)))
#else
#define SDIC_FAST_ATTR
#endif
template <int PPL, bool K6, typename TD, bool COEF = false>
__global__ __launch_bounds__(256) SDIC_FAST_ATTR void k_fast(SolveArgs a) {
    const int wpb = (int)(blockDim.x >> 6);
    const int k = xcd_remap(blockIdx.x, gridDim.x) * wpb + (int)(threadIdx.x >> 6);
    if (k >= a.nitems) return;
    const int id = a.items ? a.items[k] : (a.tile > 0 && !a.tasks) ? tile_order(k, a.grid.gw, a.grid.gh, a.tile) : k;
    Task t;
    if (!item_task(a, id, t)) return;
    SolveOut o = {};
    const bool fin = fast_solve<PPL, K6, TD, COEF>(t, a.pool, a.refI, a.refGx, a.refGy, a.refGx16, a.refGy16,
                                                   a.grad_bf16 != 0, a.chol, (const TD*)a.def, a.defc, a.w, a.h,
                                                   a.dim, a.lm != 0, a.stop2, o);
    if ((threadIdx.x & 63) != 0) return;
    if (!fin) {
        a.bail[atomicAdd(a.nbail, 1)] = id;
        return;
    }
    commit(a, id, o);
}

// Full solver over the bail list: a fixed grid of NW-wave workgroups, each
// striding over the list (its length is read on the device, so no host round
// trip). Two waves per SIMD up to 32 pixels per lane (a list of a few hundred
// subsets is solved in one round); the two largest classes take one, with
// AGPRs as spill room instead of scratch.
template <int PPL, bool K6, typename TD, int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(PPL <= 32 ? 2 : 1))) void k_slow(SolveArgs a) {
    // each wave keeps the solver's uniform state in its own copy; the waves
    // meet only in the group sums
    __shared__ WaveLds Ls[NW];
    __shared__ float red[2 * NW * kGroupSlots + 40];   // group-sum halves, then the 6x6 exchange
    WaveLds& L = Ls[threadIdx.x >> 6];
    const int nb = *a.nbail;
    for (int k = blockIdx.x; k < nb; k += gridDim.x) {
        const int id = a.bail[k];
        Task t;
        if (!item_task(a, id, t)) continue;
        SolveOut o = {};
        full_solve<PPL, K6, TD, NW>(t, a.pool, a.refI, a.refGx, a.refGy, a.refGx16, a.refGy16, a.grad_bf16 != 0, a.chol,
                                    (const TD*)a.def, a.w, a.h, a.dim, a.lm != 0, a.tol, L, red, o);
        if (threadIdx.x == 0) commit(a, id, o);
        __syncthreads();
    }
}

}  // namespace sdic
