// Wavefront-level helpers for CDNA4 (64 lanes): butterfly reductions whose
// result is bit-identical in every lane (each butterfly step adds the same two
// operands in every lane pair, and float addition is commutative), and a
// small 6x6 / 3x3 algebra kit evaluated redundantly by all lanes on
// wave-uniform values.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sdic {

constexpr int kWave = 64;

// Cross-lane moves without LDS: DPP within rows of 16 (quad_perm [1,0,3,2],
// quad_perm [2,3,0,1], row_half_mirror, row_mirror), then gfx950's
// v_permlane16_swap / v_permlane32_swap across rows and halves. After the two
// quad steps every lane of a quad holds the quad's sum, so the mirrors pair
// equal partial sums exactly as xor 4 and xor 8 would; each swap step adds
// (low half + high half) in the same operand order in both halves. The sum is
// therefore the same bits in all 64 lanes.
template <int Ctrl>
__device__ __forceinline__ uint32_t dpp_u(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, Ctrl, 0xF, 0xF, false);
}

__device__ __forceinline__ float wave_sum(float v) {
    v += __uint_as_float(dpp_u<0xB1>(__float_as_uint(v)));    // quad_perm [1,0,3,2]
    v += __uint_as_float(dpp_u<0x4E>(__float_as_uint(v)));    // quad_perm [2,3,0,1]
    v += __uint_as_float(dpp_u<0x141>(__float_as_uint(v)));   // row_half_mirror
    v += __uint_as_float(dpp_u<0x140>(__float_as_uint(v)));   // row_mirror
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    v = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    const auto h = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(h[0]) + __uint_as_float(h[1]);
}

// Six sums at once, each the same bits as wave_sum of that value: the four
// row steps per value, then the swap steps on pairs of values. A
// v_permlane16_swap of values a and b leaves the row pairs (0, 1) and (2, 3)
// of a in rows 0 and 2 and those of b in rows 1 and 3, so one swap and one add
// serve two values; a v_permlane32_swap of two such results leaves the totals
// of four values in rows 0 .. 3. Each total is read from a lane that holds it.
__device__ __forceinline__ void wave_sum6(float (&v)[6]) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        float x = v[k];
        x += __uint_as_float(dpp_u<0xB1>(__float_as_uint(x)));
        x += __uint_as_float(dpp_u<0x4E>(__float_as_uint(x)));
        x += __uint_as_float(dpp_u<0x141>(__float_as_uint(x)));
        x += __uint_as_float(dpp_u<0x140>(__float_as_uint(x)));
        v[k] = x;
    }
    float s[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v[2 * i]), __float_as_uint(v[2 * i + 1]),
                                                        false, false);
        s[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    const auto h = __builtin_amdgcn_permlane32_swap(__float_as_uint(s[0]), __float_as_uint(s[1]), false, false);
    const uint32_t t = __float_as_uint(__uint_as_float(h[0]) + __uint_as_float(h[1]));
    const auto g = __builtin_amdgcn_permlane32_swap(__float_as_uint(s[2]), __float_as_uint(s[2]), false, false);
    const uint32_t u = __float_as_uint(__uint_as_float(g[0]) + __uint_as_float(g[1]));
    v[0] = __uint_as_float((uint32_t)__builtin_amdgcn_readlane((int)t, 0));
    v[1] = __uint_as_float((uint32_t)__builtin_amdgcn_readlane((int)t, 16));
    v[2] = __uint_as_float((uint32_t)__builtin_amdgcn_readlane((int)t, 32));
    v[3] = __uint_as_float((uint32_t)__builtin_amdgcn_readlane((int)t, 48));
    v[4] = __uint_as_float((uint32_t)__builtin_amdgcn_readlane((int)u, 0));
    v[5] = __uint_as_float((uint32_t)__builtin_amdgcn_readlane((int)u, 16));
}

__device__ __forceinline__ int wave_sum_i(int v) {
    v += (int)dpp_u<0xB1>((uint32_t)v);
    v += (int)dpp_u<0x4E>((uint32_t)v);
    v += (int)dpp_u<0x141>((uint32_t)v);
    v += (int)dpp_u<0x140>((uint32_t)v);
    const auto r = __builtin_amdgcn_permlane16_swap((uint32_t)v, (uint32_t)v, false, false);
    v = (int)r[0] + (int)r[1];
    const auto h = __builtin_amdgcn_permlane32_swap((uint32_t)v, (uint32_t)v, false, false);
    return (int)h[0] + (int)h[1];
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

// A 16-bit stored value (a gradient plane's copy) back to float: fp16 bits,
// or bf16 bits (the upper half of an fp32).
__device__ __forceinline__ float half_bits_to_float(uint16_t v, bool bf16) {
    return bf16 ? __uint_as_float((uint32_t)v << 16) : (float)__builtin_bit_cast(_Float16, v);
}

__device__ __forceinline__ bool finite_f(float f) {
    return (__float_as_uint(f) & 0x7f800000u) != 0x7f800000u;
}

// 3x3 inverse by cofactors, determinant from the first column.
__device__ inline void inv3(const float* a, float* r) {
    auto cof = [&](int i, int j) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        return a[i1 * 3 + j1] * a[i2 * 3 + j2] - a[i1 * 3 + j2] * a[i2 * 3 + j1];
    };
    const float det = cof(0, 0) * a[0] + cof(1, 0) * a[3] + cof(2, 0) * a[6];
    const float inv = 1.0f / det;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r[i * 3 + j] = cof(j, i) * inv;
}

// W * inv3(A) for affine 3x3 matrices (last rows 0 0 1), with the products
// inv3 and mul3 form for the entries that are not multiplied by an exact zero
// of those last rows, in their order: the same bits as mul3(W, inv3(A)) for
// finite entries, up to the sign of an entry that is zero.
__device__ __forceinline__ void compose_inverse_affine(const float* W, const float* a, float* r) {
    const float det = a[4] * a[0] + -a[1] * a[3];
    const float inv = 1.0f / det;
    // inv3's r[0..5] and r[8] (r[6] = r[7] = 0)
    const float i0 = a[4] * inv, i1 = -a[1] * inv, i2 = (a[1] * a[5] - a[2] * a[4]) * inv;
    const float i3 = -a[3] * inv, i4 = a[0] * inv, i5 = (a[2] * a[3] - a[0] * a[5]) * inv;
    const float i8 = det * inv;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const float w0 = W[i * 3], w1 = W[i * 3 + 1], w2 = W[i * 3 + 2];
        r[i * 3] = w0 * i0 + w1 * i3;
        r[i * 3 + 1] = w0 * i1 + w1 * i4;
        r[i * 3 + 2] = w0 * i2 + w1 * i5 + w2 * i8;
    }
    r[6] = 0.0f;
    r[7] = 0.0f;
    r[8] = W[8] * i8;
}

__device__ inline void mul3(const float* a, const float* b, float* r) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            float s = 0.0f;
#pragma unroll
            for (int k = 0; k < 3; ++k) s += a[i * 3 + k] * b[k * 3 + j];
            r[i * 3 + j] = s;
        }
}

// Gauss-Jordan inverse of a 6x6 with partial pivoting -- the CPU oracle's
// sequence -- on a matrix held in registers (every index is a compile-time
// constant once unrolled; the pivot row moves by selects): each element sees
// the oracle's operations in its order.
__device__ __forceinline__ void gj_inverse_reg(const float* A, float* out) {
    float m[6][12];
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            m[r][c] = A[r * 6 + c];
            m[r][6 + c] = (r == c) ? 1.0f : 0.0f;
        }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        int piv = k;
        float best = fabsf(m[k][k]);
#pragma unroll
        for (int r = k + 1; r < 6; ++r)
            if (fabsf(m[r][k]) > best) {
                best = fabsf(m[r][k]);
                piv = r;
            }
#pragma unroll
        for (int r = k + 1; r < 6; ++r) {
            const bool sw = piv == r;
#pragma unroll
            for (int c = 0; c < 12; ++c) {
                const float t = m[k][c];
                m[k][c] = sw ? m[r][c] : t;
                m[r][c] = sw ? t : m[r][c];
            }
        }
        const float ip = 1.0f / m[k][k];
#pragma unroll
        for (int c = 0; c < 12; ++c) m[k][c] *= ip;
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            if (r == k) continue;
            const float f = m[r][k];
#pragma unroll
            for (int c = 0; c < 12; ++c) m[r][c] = (f == 0.0f) ? m[r][c] : m[r][c] - f * m[k][c];
        }
    }
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) out[r * 6 + c] = m[r][6 + c];
}

__device__ __forceinline__ float norm1_reg(const float* m) {
    float best = 0.0f;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        float s = 0.0f;
#pragma unroll
        for (int r = 0; r < 6; ++r) s += fabsf(m[r * 6 + c]);
        best = s > best ? s : best;
    }
    return best;
}

// hessian::well_conditioned (hessian.hpp:40-52) in registers.
__device__ __forceinline__ bool well_conditioned_reg(const float* H, int half) {
    const float s = half > 1 ? 1.0f / (float)half : 1.0f;
    float Hn[36], Hi[36];
    bool fin = true;
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const float dr = r < 2 ? 1.0f : s, dc = c < 2 ? 1.0f : s;
            Hn[r * 6 + c] = dr * H[r * 6 + c] * dc;
            fin = fin && finite_f(Hn[r * 6 + c]);
        }
    if (!fin) return false;
    gj_inverse_reg(Hn, Hi);
#pragma unroll
    for (int q = 0; q < 36; ++q) fin = fin && finite_f(Hi[q]);
    if (!fin) return false;
    const float a = norm1_reg(Hn), b = norm1_reg(Hi);
    return 1.0f / (a * b) > 1e-6f;
}

// ---- the same two routines with the matrix spread over a wave --------------
// Lane l = 6 r + c (l < 36) holds entry (r, c) of a 6x6 in `mL` and, for the
// Gauss-Jordan sweep, entry (r, 6 + c) of the augmented half in `mR`; lanes
// 36..63 carry values nobody reads. Each entry sees exactly the operations of
// the register routines above, in their order, so the results are theirs bit
// for bit; a pivot step costs a few cross-lane moves instead of a sweep over
// 72 registers in every lane.
__device__ __forceinline__ float lane_read(float v, int l) {
    return __uint_as_float((uint32_t)__builtin_amdgcn_readlane((int)__float_as_uint(v), l));
}
__device__ __forceinline__ float lane_fetch(float v, int src) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
}

// gj_inverse_reg on (mL | mR) = (A | I): the inverse ends in mR.
__device__ __forceinline__ void gj_inverse_lanes(float& mL, float& mR, int lane) {
    const int r = lane / 6, c = lane - 6 * r;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        int piv = k;
        float best = fabsf(lane_read(mL, 6 * k + k));
#pragma unroll
        for (int q = k + 1; q < 6; ++q) {
            const float v = fabsf(lane_read(mL, 6 * q + k));
            if (v > best) {
                best = v;
                piv = q;
            }
        }
        if (piv != k) {
            const int src = (r == k) ? 6 * piv + c : (r == piv) ? 6 * k + c : lane;
            mL = lane_fetch(mL, src);
            mR = lane_fetch(mR, src);
        }
        const float ip = 1.0f / lane_read(mL, 6 * k + k);
        if (r == k) {
            mL *= ip;
            mR *= ip;
        }
        // this row's factor m[r][k] and the pivot row's entries of this column
        const float f = lane_fetch(mL, 6 * r + k);
        const float kL = lane_fetch(mL, 6 * k + c), kR = lane_fetch(mR, 6 * k + c);
        if (r != k) {
            mL = (f == 0.0f) ? mL : mL - f * kL;
            mR = (f == 0.0f) ? mR : mR - f * kR;
        }
    }
}

// norm1_reg of the matrix in v: column sums on lanes 0..5 (rows in order),
// then their maximum in column order (uniform).
__device__ __forceinline__ float norm1_lanes(float v, int lane) {
    const int c = lane < 6 ? lane : 0;
    float s = 0.0f;
#pragma unroll
    for (int r = 0; r < 6; ++r) s += fabsf(lane_fetch(v, 6 * r + c));
    float best = 0.0f;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const float sq = lane_read(s, q);
        best = sq > best ? sq : best;
    }
    return best;
}

// well_conditioned_reg of the matrix in hL.
__device__ __forceinline__ bool well_conditioned_lanes(float hL, int half, int lane) {
    const int r = lane / 6, c = lane - 6 * r;
    const float s = half > 1 ? 1.0f / (float)half : 1.0f;
    const float dr = r < 2 ? 1.0f : s, dc = c < 2 ? 1.0f : s;
    const float hn = dr * hL * dc;
    if (__builtin_amdgcn_ballot_w64(lane < 36 && !finite_f(hn)) != 0) return false;
    float mL = hn, mR = (r == c) ? 1.0f : 0.0f;
    gj_inverse_lanes(mL, mR, lane);
    if (__builtin_amdgcn_ballot_w64(lane < 36 && !finite_f(mR)) != 0) return false;
    const float a = norm1_lanes(hn, lane), b = norm1_lanes(mR, lane);
    return 1.0f / (a * b) > 1e-6f;
}

// FP64 Cholesky factor of a symmetric positive definite 6x6 (float input,
// converted exactly): L row-major lower triangle at i (i + 1) / 2 + j, its
// diagonal stored as the reciprocal. False when a pivot is not positive.
__device__ __forceinline__ bool chol64_factor(const float* A, double* Lp) {
    double L[6][6];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            double s = (double)A[i * 6 + j];
#pragma unroll
            for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
            if (i == j) {
                if (!(s > 0.0)) return false;
                L[i][i] = sqrt(s);
            } else {
                L[i][j] = s / L[j][j];
            }
        }
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) Lp[i * (i + 1) / 2 + j] = (i == j) ? 1.0 / L[i][i] : L[i][j];
    return true;
}

// The Gauss-Newton step d = -(L L^T)^-1 g by FP64 forward and back
// substitution on a factor from chol64_factor.
__device__ __forceinline__ void chol64_step(const double* __restrict__ Lp, const float* g, float* d) {
    double y[6], x[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double s = (double)g[i];
#pragma unroll
        for (int k = 0; k < i; ++k) s -= Lp[i * (i + 1) / 2 + k] * y[k];
        y[i] = s * Lp[i * (i + 1) / 2 + i];
    }
#pragma unroll
    for (int i = 5; i >= 0; --i) {
        double s = y[i];
#pragma unroll
        for (int k = i + 1; k < 6; ++k) s -= Lp[k * (k + 1) / 2 + i] * x[k];
        x[i] = s * Lp[i * (i + 1) / 2 + i];
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) d[i] = -(float)x[i];
}

// blockIdx -> work index so that consecutive work items land on one XCD:
// the dispatcher deals workgroups round-robin over the 8 XCDs, each with its
// own L2, and neighbouring subsets share most of their pixels.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
    constexpr int kXcd = 8;
    const int per = (nblocks + kXcd - 1) / kXcd;
    const int xcd = bid % kXcd, local = bid / kXcd;
    const int full = nblocks % kXcd == 0 ? kXcd : nblocks % kXcd;  // XCDs with `per` items
    // XCDs [0, full) own `per` items each, the rest `per - 1`.
    const int base = xcd < full ? xcd * per : full * per + (xcd - full) * (per - 1);
    return base + local;
}

}  // namespace sdic
