// The learned seeder's kernels and host driver (seeder.hpp): training pairs
// rendered on the GPU, the network's forward and backward passes, the loss,
// Adam, and batched inference on patches cut from the engine's level-2
// pyramids. Every reduction is a fixed-order loop or a fixed tree, so a pass
// is deterministic.
#include "seeder.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>

namespace sdic {

namespace {

using S = SeederShape;

struct HipFailS {
    hipError_t e;
};
inline void hck(hipError_t e) {
    if (e != hipSuccess) throw HipFailS{e};
}

__host__ __device__ inline uint64_t smix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// uniform [0, 1) from (seed, a, b)
__host__ __device__ inline float rnd(uint64_t seed, uint64_t a, uint64_t b) {
    return (float)(smix(seed ^ smix(a * 0x100000001B3ull + b)) >> 40) * (1.0f / 16777216.0f);
}
__device__ inline float gauss(uint64_t seed, uint64_t a, uint64_t b) {
    const float u1 = fmaxf(rnd(seed, a, 2 * b), 1e-7f), u2 = rnd(seed, a, 2 * b + 1);
    return sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2);
}

// Keys (a = -0.5) cubic weights
__device__ inline void keys4(float t, float w[4]) {
    const float t2 = t * t, t3 = t2 * t;
    w[0] = -0.5f * t3 + t2 - 0.5f * t;
    w[1] = 1.5f * t3 - 2.5f * t2 + 1.0f;
    w[2] = -1.5f * t3 + 2.0f * t2 + 0.5f * t;
    w[3] = 0.5f * t3 - 0.5f * t2;
}
__device__ inline float bicubic_u8(const uint8_t* __restrict__ img, int w, int h, float x, float y) {
    const float fx = floorf(x), fy = floorf(y);
    float wx[4], wy[4];
    keys4(x - fx, wx);
    keys4(y - fy, wy);
    const int ix = (int)fx, iy = (int)fy;
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int yy = min(max(iy - 1 + j, 0), h - 1);
        float r = 0.0f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int xx = min(max(ix - 1 + i, 0), w - 1);
            r += wx[i] * (float)img[(size_t)yy * w + xx];
        }
        s += wy[j] * r;
    }
    return s;
}

// One training pair per sample n: a random level-2 centre c2, a random affine
// warp about the patch centre P = 4 c2 - 0.5 (full resolution) and, on a
// `het_frac` share of the pairs, a smooth heterogeneous field on top of it --
// one Gaussian bump per displacement component, of amplitude up to
// `bump_amp` px and width `bump_min` .. `bump_max` px, centred within one
// width of P -- the reference patch as the engine's pyramid makes it (two
// 2x2 box averages) and the deformed patch from the warped, noisy,
// u8-rounded full-resolution samples. The target is the field's first-order
// warp at P (the affine's plus the bumps' values and gradients there),
// scaled (u / kUvScale, gradients / kGradScale).
struct PairBumps {
    float het_frac, amp, wmin, wmax;
};
__global__ void k_pairs(const uint8_t* __restrict__ ref, int w, int h, int B, uint64_t seed, uint64_t step,
                        float max_shift, float max_grad, float noise, PairBumps bm, float* __restrict__ x,
                        float* __restrict__ t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * S::kPatch * S::kPatch) return;
    const int n = i / (S::kPatch * S::kPatch), px = i % S::kPatch, py = (i / S::kPatch) % S::kPatch;
    const uint64_t id = step * 4096ull + (uint64_t)n;
    const int m2 = 64, w2 = w / 4, h2 = h / 4;
    const int c2x = m2 + (int)(rnd(seed, id, 0) * (float)(w2 - 2 * m2));
    const int c2y = m2 + (int)(rnd(seed, id, 1) * (float)(h2 - 2 * m2));
    const float tu = (2.0f * rnd(seed, id, 2) - 1.0f) * max_shift, tv = (2.0f * rnd(seed, id, 3) - 1.0f) * max_shift;
    const float g00 = (2.0f * rnd(seed, id, 4) - 1.0f) * max_grad, g01 = (2.0f * rnd(seed, id, 5) - 1.0f) * max_grad;
    const float g10 = (2.0f * rnd(seed, id, 6) - 1.0f) * max_grad, g11 = (2.0f * rnd(seed, id, 7) - 1.0f) * max_grad;
    // x = P + A (X - P) + t with A = I + G; X = P + A^-1 (x - P - t)
    const float a00 = 1.0f + g00, a01 = g01, a10 = g10, a11 = 1.0f + g11;
    const float det = a00 * a11 - a01 * a10;
    const float i00 = a11 / det, i01 = -a01 / det, i10 = -a10 / det, i11 = a00 / det;
    const float Px = 4.0f * c2x - 0.5f, Py = 4.0f * c2y - 0.5f;
    const int fx0 = 4 * (c2x - S::kPatch / 2 + px), fy0 = 4 * (c2y - S::kPatch / 2 + py);
    // the bumps: h_u(X) = au g(X - Cu), h_v(X) = av g(X - Cv), g = exp(-|.|^2 / 2 s^2)
    const bool het = rnd(seed, id, 8) < bm.het_frac;
    const float au = het ? (2.0f * rnd(seed, id, 9) - 1.0f) * bm.amp : 0.0f;
    const float av = het ? (2.0f * rnd(seed, id, 10) - 1.0f) * bm.amp : 0.0f;
    const float su = bm.wmin + rnd(seed, id, 11) * (bm.wmax - bm.wmin);
    const float sv = bm.wmin + rnd(seed, id, 12) * (bm.wmax - bm.wmin);
    const float cux = Px + (2.0f * rnd(seed, id, 13) - 1.0f) * su, cuy = Py + (2.0f * rnd(seed, id, 14) - 1.0f) * su;
    const float cvx = Px + (2.0f * rnd(seed, id, 15) - 1.0f) * sv, cvy = Py + (2.0f * rnd(seed, id, 16) - 1.0f) * sv;
    const auto bump = [&](float a, float sg, float cx, float cy, float X, float Y) {
        const float ex = X - cx, ey = Y - cy;
        return a * expf(-(ex * ex + ey * ey) / (2.0f * sg * sg));
    };
    // reference: the pyramid's two box levels
    float l1[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int bx = fx0 + 2 * (q & 1), by = fy0 + 2 * (q >> 1);
        const uint8_t* a = ref + (size_t)by * w + bx;
        l1[q] = 0.25f * (((float)a[0] + (float)a[1]) + ((float)a[w] + (float)a[w + 1]));
    }
    const float rv = 0.25f * ((l1[0] + l1[1]) + (l1[2] + l1[3]));
    // deformed: warped samples, noise, u8 rounding, then the same two levels
    float d1[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float s4[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int xx = fx0 + 2 * (q & 1) + (r & 1), yy = fy0 + 2 * (q >> 1) + (r >> 1);
            // X with x = X + t + G (X - P) + h(X): the affine inverse, then
            // fixed-point steps on the bumps (a contraction: their gradients
            // stay far below one)
            float X = 0.0f, Y = 0.0f, hu = 0.0f, hv = 0.0f;
#pragma unroll 1
            for (int itx = 0; itx < (het ? 6 : 1); ++itx) {
                const float dx = (float)xx - Px - tu - hu, dy = (float)yy - Py - tv - hv;
                X = Px + i00 * dx + i01 * dy;
                Y = Py + i10 * dx + i11 * dy;
                if (het) {
                    hu = bump(au, su, cux, cuy, X, Y);
                    hv = bump(av, sv, cvx, cvy, X, Y);
                }
            }
            float v = bicubic_u8(ref, w, h, X, Y) + noise * gauss(seed ^ 0x5EEDull, id, (uint64_t)(yy * 8191 + xx));
            v = fminf(fmaxf(rintf(v), 0.0f), 255.0f);
            s4[r] = v;
        }
        d1[q] = 0.25f * ((s4[0] + s4[1]) + (s4[2] + s4[3]));
    }
    const float dv = 0.25f * ((d1[0] + d1[1]) + (d1[2] + d1[3]));
    const size_t base = (size_t)n * 2 * S::kPatch * S::kPatch + (size_t)py * S::kPatch + px;
    x[base] = rv;
    x[base + S::kPatch * S::kPatch] = dv;
    if (px == 0 && py == 0) {
        // the field's value and gradient at P
        const float hu0 = bump(au, su, cux, cuy, Px, Py), hv0 = bump(av, sv, cvx, cvy, Px, Py);
        const float hux = -hu0 * (Px - cux) / (su * su), huy = -hu0 * (Py - cuy) / (su * su);
        const float hvx = -hv0 * (Px - cvx) / (sv * sv), hvy = -hv0 * (Py - cvy) / (sv * sv);
        float* tt = t + (size_t)n * S::kOut;
        tt[0] = (tu + hu0) / S::kUvScale;
        tt[1] = (tv + hv0) / S::kUvScale;
        tt[2] = (g00 + hux) / S::kGradScale;
        tt[3] = (g01 + huy) / S::kGradScale;
        tt[4] = (g10 + hvx) / S::kGradScale;
        tt[5] = (g11 + hvy) / S::kGradScale;
    }
}

// Patch pairs for inference from the level-2 planes.
__global__ void k_cut(const float* __restrict__ r2, const float* __restrict__ d2, int w2, int h2,
                      const int* __restrict__ c, const int* __restrict__ off, int n, float* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * S::kPatch * S::kPatch) return;
    const int k = i / (S::kPatch * S::kPatch), px = i % S::kPatch, py = (i / S::kPatch) % S::kPatch;
    const int rx = min(max(c[2 * k] - S::kPatch / 2 + px, 0), w2 - 1);
    const int ry = min(max(c[2 * k + 1] - S::kPatch / 2 + py, 0), h2 - 1);
    const int dx = min(max(rx + off[2 * k], 0), w2 - 1), dy = min(max(ry + off[2 * k + 1], 0), h2 - 1);
    const size_t base = (size_t)k * 2 * S::kPatch * S::kPatch + (size_t)py * S::kPatch + px;
    x[base] = r2[(size_t)ry * w2 + rx];
    x[base + S::kPatch * S::kPatch] = d2[(size_t)dy * w2 + dx];
}

// Zero mean, unit variance per (sample, channel) plane: one block of 256.
__global__ __launch_bounds__(256) void k_normalize(float* __restrict__ x) {
    constexpr int N = S::kPatch * S::kPatch;
    float* p = x + (size_t)blockIdx.x * N;
    __shared__ float red[256];
    float s = 0.0f;
    for (int i = threadIdx.x; i < N; i += 256) s += p[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const float mean = red[0] / (float)N;
    __syncthreads();
    float q = 0.0f;
    for (int i = threadIdx.x; i < N; i += 256) {
        const float d = p[i] - mean;
        q += d * d;
    }
    red[threadIdx.x] = q;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const float inv = 1.0f / sqrtf(red[0] / (float)N + 1e-6f);
    for (int i = threadIdx.x; i < N; i += 256) p[i] = (p[i] - mean) * inv;
}

// 3x3 convolution, stride 2, padding 1, then ReLU. out: B x 32 x hin/2 x hin/2.
__global__ void k_conv_fwd(const float* __restrict__ in, int cin, int hin, const float* __restrict__ wt,
                           const float* __restrict__ bias, float* __restrict__ out, int B) {
    const int ho = hin / 2;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * S::kCh * ho * ho) return;
    const int x = i % ho, y = (i / ho) % ho, co = (i / (ho * ho)) % S::kCh, n = i / (S::kCh * ho * ho);
    float s = bias[co];
    for (int ci = 0; ci < cin; ++ci) {
        const float* ip = in + ((size_t)n * cin + ci) * hin * hin;
        const float* wp = wt + ((size_t)co * cin + ci) * 9;
#pragma unroll
        for (int ky = 0; ky < 3; ++ky) {
            const int iy = 2 * y + ky - 1;
            if (iy < 0 || iy >= hin) continue;
#pragma unroll
            for (int kx = 0; kx < 3; ++kx) {
                const int ix = 2 * x + kx - 1;
                if (ix < 0 || ix >= hin) continue;
                s += wp[ky * 3 + kx] * ip[iy * hin + ix];
            }
        }
    }
    out[i] = s > 0.0f ? s : 0.0f;
}

// gradient of the input of a convolution from the gradient of its
// pre-activation output (g), then masked by the input's own ReLU (a_in > 0)
__global__ void k_conv_bwd_in(const float* __restrict__ g, int cin, int hin, const float* __restrict__ wt,
                              const float* __restrict__ a_in, float* __restrict__ gin, int B) {
    const int ho = hin / 2;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * cin * hin * hin) return;
    const int ix = i % hin, iy = (i / hin) % hin, ci = (i / (hin * hin)) % cin, n = i / (cin * hin * hin);
    if (!(a_in[i] > 0.0f)) {
        gin[i] = 0.0f;
        return;
    }
    float s = 0.0f;
    for (int co = 0; co < S::kCh; ++co) {
        const float* gp = g + ((size_t)n * S::kCh + co) * ho * ho;
        const float* wp = wt + ((size_t)co * cin + ci) * 9;
#pragma unroll
        for (int ky = 0; ky < 3; ++ky) {
            const int t = iy + 1 - ky;
            if (t < 0 || (t & 1) || (t >> 1) >= ho) continue;
            const int y = t >> 1;
#pragma unroll
            for (int kx = 0; kx < 3; ++kx) {
                const int u = ix + 1 - kx;
                if (u < 0 || (u & 1) || (u >> 1) >= ho) continue;
                s += wp[ky * 3 + kx] * gp[y * ho + (u >> 1)];
            }
        }
    }
    gin[i] = s;
}

// weight and bias gradients of a convolution: one block per weight (then one
// per bias), a fixed tree over the block
__global__ __launch_bounds__(256) void k_conv_bwd_w(const float* __restrict__ in, int cin, int hin,
                                                    const float* __restrict__ g, float* __restrict__ gw,
                                                    float* __restrict__ gb, int B) {
    const int ho = hin / 2, nw = S::kCh * cin * 9;
    const int k = blockIdx.x;
    const int m = B * ho * ho;
    float s = 0.0f;
    if (k < nw) {
        const int kx = k % 3, ky = (k / 3) % 3, ci = (k / 9) % cin, co = k / (9 * cin);
        for (int j = threadIdx.x; j < m; j += 256) {
            const int x = j % ho, y = (j / ho) % ho, n = j / (ho * ho);
            const int iy = 2 * y + ky - 1, ix = 2 * x + kx - 1;
            if (iy < 0 || iy >= hin || ix < 0 || ix >= hin) continue;
            s += g[(((size_t)n * S::kCh + co) * ho + y) * ho + x] * in[(((size_t)n * cin + ci) * hin + iy) * hin + ix];
        }
    } else {
        const int co = k - nw;
        for (int j = threadIdx.x; j < m; j += 256) {
            const int x = j % ho, y = (j / ho) % ho, n = j / (ho * ho);
            s += g[(((size_t)n * S::kCh + co) * ho + y) * ho + x];
        }
    }
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (k < nw) gw[k] = red[0];
        else gb[k - nw] = red[0];
    }
}

__global__ void k_fc_fwd(const float* __restrict__ f, const float* __restrict__ wt, const float* __restrict__ b,
                         float* __restrict__ y, int B) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * S::kOut) return;
    const int k = i % S::kOut, n = i / S::kOut;
    float s = b[k];
    for (int j = 0; j < S::kFeat; ++j) s += wt[k * S::kFeat + j] * f[(size_t)n * S::kFeat + j];
    y[i] = s;
}

// mean squared error over B x 6: gradient and per-entry squared error
__global__ void k_loss(const float* __restrict__ y, const float* __restrict__ t, float* __restrict__ gy,
                       float* __restrict__ sq, int B) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * S::kOut) return;
    const float d = y[i] - t[i];
    gy[i] = 2.0f * d / (float)(B * S::kOut);
    sq[i] = d * d;
}

__global__ void k_fc_bwd_w(const float* __restrict__ f, const float* __restrict__ gy, float* __restrict__ gw,
                           float* __restrict__ gb, int B) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S::kOut * S::kFeat + S::kOut) return;
    float s = 0.0f;
    if (i < S::kOut * S::kFeat) {
        const int k = i / S::kFeat, j = i % S::kFeat;
        for (int n = 0; n < B; ++n) s += gy[n * S::kOut + k] * f[(size_t)n * S::kFeat + j];
        gw[i] = s;
    } else {
        const int k = i - S::kOut * S::kFeat;
        for (int n = 0; n < B; ++n) s += gy[n * S::kOut + k];
        gb[k] = s;
    }
}

__global__ void k_fc_bwd_in(const float* __restrict__ gy, const float* __restrict__ wt, const float* __restrict__ f,
                            float* __restrict__ gf, int B) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * S::kFeat) return;
    const int j = i % S::kFeat, n = i / S::kFeat;
    if (!(f[i] > 0.0f)) {
        gf[i] = 0.0f;
        return;
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < S::kOut; ++k) s += wt[k * S::kFeat + j] * gy[n * S::kOut + k];
    gf[i] = s;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float bc1, float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi, vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

inline int blocks(size_t n, int bs) { return (int)((n + bs - 1) / bs); }

}  // namespace

namespace {

// ---- inference on MFMA ------------------------------------------------------
// One workgroup of 8 waves per patch pair. Activations stay in LDS as fp16,
// channel-last with kPS halves per position (32 channels and 8 of padding);
// each 3x3 stride-2 convolution is an implicit GEMM on
// v_mfma_f32_16x16x16f16 (FP32 accumulation): rows are 16 output positions,
// columns 16 output channels, K runs over (tap, input channel) with the
// channel fastest, so a lane's four K values are four adjacent channels of
// one input position, one 8-byte LDS read. Wave w takes channel tile w / 4
// and position tiles w % 4, w % 4 + 4, ...; its B fragments (the layer's
// fp16 weights for its 16 channels) are held in registers for the layer.
// Operand layout of the MFMA: lane l gives A[l % 16][4 (l / 16) + j] and
// B[4 (l / 16) + j][l % 16]; D[r][c] returns in lane 16 (r / 4) + c, slot
// r % 4.
typedef _Float16 sd_half4 __attribute__((ext_vector_type(4)));
typedef _Float16 sd_half2 __attribute__((ext_vector_type(2)));
typedef float sd_float4 __attribute__((ext_vector_type(4)));

namespace mf {
constexpr int kPS = 40;                     // halves per position
constexpr int kIn = 0;                      // input planes, 64 x 64 x 2
constexpr int kA1 = 64 * 64 * 2;            // conv 1 out, 32 x 32 x kPS
constexpr int kA2 = kA1 + 32 * 32 * kPS;    // conv 2 out, 16 x 16 x kPS
constexpr int kA3 = kA1;                    // conv 3 out, 8 x 8 x kPS (over conv 1's)
constexpr int kHalves = kA2 + 16 * 16 * kPS;
constexpr int kW1 = 32 * 32;                // conv 1 [n][k = tap * 2 + ci], 18 of 32 used
constexpr int kWl = 32 * 288;               // conv 2..4 [n][tap * 32 + ci]
constexpr int kHalfWeights = kW1 + 3 * kWl;
}  // namespace mf

// FP32 parameters -> the fp16 weight layouts of the MFMA kernel.
__global__ void k_half_weights(const float* __restrict__ P, _Float16* __restrict__ wh) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= mf::kHalfWeights) return;
    float v = 0.0f;
    if (i < mf::kW1) {
        const int n = i / 32, k = i % 32;
        if (k < 18) v = P[S::off_w(0) + (n * S::kInCh + (k & 1)) * 9 + (k >> 1)];
    } else {
        const int l = 1 + (i - mf::kW1) / mf::kWl, r = (i - mf::kW1) % mf::kWl;
        const int n = r / 288, kk = r % 288;
        v = P[S::off_w(l) + (n * S::kCh + (kk & 31)) * 9 + (kk >> 5)];
    }
    wh[i] = (_Float16)v;
}

__device__ inline float wave_sum64(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// conv 2..4: in (HIN x HIN x kPS) -> out (HIN / 2 square x kPS), or, for the
// last layer, FP32 features in the network's C x H x W order.
template <int HIN, bool LAST>
__device__ inline void mf_conv(const _Float16* __restrict__ in, _Float16* __restrict__ out, float* __restrict__ feat,
                               const _Float16* __restrict__ wl, const float* __restrict__ bias, int wv, int lane) {
    constexpr int HO = HIN / 2, M = HO * HO, TILES = M / 16;
    const int nt = wv >> 2, g = lane >> 4, col = lane & 15, n = nt * 16 + col;
    if ((wv & 3) >= TILES) return;
    sd_half4 b[18];
#pragma unroll
    for (int c = 0; c < 18; ++c) b[c] = *reinterpret_cast<const sd_half4*>(wl + n * 288 + 16 * c + 4 * g);
    const float bn = bias[n];
    for (int pt = wv & 3; pt < TILES; pt += 4) {
        const int m = pt * 16 + col, oy = m / HO, ox = m % HO;
        sd_float4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int c = 0; c < 18; ++c) {
            const int tap = c >> 1, ky = tap / 3, kx = tap % 3;
            const int iy = 2 * oy + ky - 1, ix = 2 * ox + kx - 1;
            sd_half4 a = {(_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f};
            if (iy >= 0 && iy < HIN && ix >= 0 && ix < HIN)
                a = *reinterpret_cast<const sd_half4*>(in + (iy * HIN + ix) * mf::kPS + 16 * (c & 1) + 4 * g);
            acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a, b[c], acc, 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int mm = pt * 16 + 4 * g + j;
            const float v = fmaxf(acc[j] + bn, 0.0f);
            if (LAST) feat[n * M + mm] = v;
            else out[mm * mf::kPS + n] = (_Float16)v;
        }
    }
}

// CUT: cut the pair from the level-2 planes (as k_cut) and normalise each
// plane (as k_normalize); otherwise read pre-normalised pairs from xin.
// y: B x 6 outputs in the network's scaled units.
template <bool CUT>
__global__ __launch_bounds__(512) void k_seed_mfma(const float* __restrict__ r2, const float* __restrict__ d2, int w2,
                                                   int h2, const int* __restrict__ c, const int* __restrict__ off,
                                                   const float* __restrict__ xin, const _Float16* __restrict__ wh,
                                                   const float* __restrict__ P, float* __restrict__ y) {
    __shared__ __attribute__((aligned(16))) _Float16 lds[mf::kHalves];
    __shared__ float red[16];
    const int k = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (int ch = 0; ch < 2; ++ch) {
        float v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int i = t + 512 * q, px = i & 63, py = i >> 6;
            if (CUT) {
                const int rx = min(max(c[2 * k] - S::kPatch / 2 + px, 0), w2 - 1);
                const int ry = min(max(c[2 * k + 1] - S::kPatch / 2 + py, 0), h2 - 1);
                if (ch == 0) {
                    v[q] = r2[(size_t)ry * w2 + rx];
                } else {
                    const int dx = min(max(rx + off[2 * k], 0), w2 - 1), dy = min(max(ry + off[2 * k + 1], 0), h2 - 1);
                    v[q] = d2[(size_t)dy * w2 + dx];
                }
            } else {
                v[q] = xin[((size_t)k * 2 + ch) * 4096 + i];
            }
        }
        float mean = 0.0f, inv = 1.0f;
        if (CUT) {
            float s = 0.0f;
#pragma unroll
            for (int q = 0; q < 8; ++q) s += v[q];
            s = wave_sum64(s);
            if (lane == 0) red[wv] = s;
            __syncthreads();
            float tot = 0.0f;
            for (int w = 0; w < 8; ++w) tot += red[w];
            mean = tot / 4096.0f;
            __syncthreads();
            float q2 = 0.0f;
#pragma unroll
            for (int q = 0; q < 8; ++q) q2 += (v[q] - mean) * (v[q] - mean);
            q2 = wave_sum64(q2);
            if (lane == 0) red[8 + wv] = q2;
            __syncthreads();
            float tq = 0.0f;
            for (int w = 0; w < 8; ++w) tq += red[8 + w];
            inv = 1.0f / sqrtf(tq / 4096.0f + 1e-6f);
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) lds[mf::kIn + (t + 512 * q) * 2 + ch] = (_Float16)((v[q] - mean) * inv);
    }
    __syncthreads();
    // conv 1: 64 x 64 x 2 -> 32 x 32 x 32; K = 9 taps x 2 channels (k = tap * 2 + ci), padded to 32
    {
        const int nt = wv >> 2, g = lane >> 4, col = lane & 15, n = nt * 16 + col;
        const sd_half4 b0 = *reinterpret_cast<const sd_half4*>(wh + n * 32 + 4 * g);
        const sd_half4 b1 = *reinterpret_cast<const sd_half4*>(wh + n * 32 + 16 + 4 * g);
        const float bn = P[S::off_b(0) + n];
        for (int pt = wv & 3; pt < 64; pt += 4) {
            const int oy = pt >> 1, ox = (pt & 1) * 16 + col;
            sd_float4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
                sd_half4 a;
#pragma unroll
                for (int jj = 0; jj < 2; ++jj) {
                    const int tap = 8 * cc + 2 * g + jj;
                    sd_half2 pr = {(_Float16)0.0f, (_Float16)0.0f};
                    if (tap < 9) {
                        const int iy = 2 * oy + tap / 3 - 1, ix = 2 * ox + tap % 3 - 1;
                        if (iy >= 0 && iy < 64 && ix >= 0 && ix < 64)
                            pr = *reinterpret_cast<const sd_half2*>(lds + mf::kIn + (iy * 64 + ix) * 2);
                    }
                    a[2 * jj] = pr[0];
                    a[2 * jj + 1] = pr[1];
                }
                acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a, cc == 0 ? b0 : b1, acc, 0, 0, 0);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int x = (pt & 1) * 16 + 4 * g + j;
                lds[mf::kA1 + (oy * 32 + x) * mf::kPS + n] = (_Float16)fmaxf(acc[j] + bn, 0.0f);
            }
        }
    }
    __syncthreads();
    mf_conv<32, false>(lds + mf::kA1, lds + mf::kA2, nullptr, wh + mf::kW1, P + S::off_b(1), wv, lane);
    __syncthreads();
    mf_conv<16, false>(lds + mf::kA2, lds + mf::kA3, nullptr, wh + mf::kW1 + mf::kWl, P + S::off_b(2), wv, lane);
    __syncthreads();
    float* feat = reinterpret_cast<float*>(lds + mf::kIn);   // the input planes are spent
    mf_conv<8, true>(lds + mf::kA3, nullptr, feat, wh + mf::kW1 + 2 * mf::kWl, P + S::off_b(3), wv, lane);
    __syncthreads();
    if (wv < S::kOut) {
        const float* wf = P + S::off_fw() + wv * S::kFeat;
        float s = 0.0f;
#pragma unroll
        for (int q = 0; q < S::kFeat / 64; ++q) s += wf[lane + 64 * q] * feat[lane + 64 * q];
        s = wave_sum64(s);
        if (lane == 0) y[(size_t)k * S::kOut + wv] = s + P[S::off_fb() + wv];
    }
}

}  // namespace

struct Seeder::Impl {
    hipStream_t st = nullptr;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    int cap = 0;
    float *x = nullptr, *a1 = nullptr, *a2 = nullptr, *a3 = nullptr, *a4 = nullptr, *y = nullptr, *t = nullptr;
    float *gy = nullptr, *g1 = nullptr, *g2 = nullptr, *g3 = nullptr, *g4 = nullptr, *sq = nullptr;
    int *ci = nullptr, *oi = nullptr;
    int icap = 0;
    _Float16* wh = nullptr;   // fp16 weights of the MFMA inference
    int pver = 0, hver = -1;  // parameter writes so far; the one wh was made from

    ~Impl() { release(); }
    void release() {
        float** fs[] = {&p, &g, &m, &v, &x, &a1, &a2, &a3, &a4, &y, &t, &gy, &g1, &g2, &g3, &g4, &sq};
        for (float** f : fs)
            if (*f) {
                (void)hipFree(*f);
                *f = nullptr;
            }
        if (ci) (void)hipFree(ci);
        if (oi) (void)hipFree(oi);
        if (wh) (void)hipFree(wh);
        ci = oi = nullptr;
        wh = nullptr;
        cap = icap = 0;
        hver = -1;
    }
    void params_alloc() {
        if (p) return;
        const size_t nb = sizeof(float) * S::total();
        hck(hipMalloc(&p, nb));
        hck(hipMalloc(&g, nb));
        hck(hipMalloc(&m, nb));
        hck(hipMalloc(&v, nb));
        hck(hipMemsetAsync(m, 0, nb, st));
        hck(hipMemsetAsync(v, 0, nb, st));
    }
    void batch_alloc(int B) {
        if (B <= cap) return;
        float** fs[] = {&x, &a1, &a2, &a3, &a4, &y, &t, &gy, &g1, &g2, &g3, &g4, &sq};
        for (float** f : fs)
            if (*f) {
                (void)hipFree(*f);
                *f = nullptr;
            }
        const size_t n = (size_t)B;
        hck(hipMalloc(&x, sizeof(float) * n * 2 * 64 * 64));
        hck(hipMalloc(&a1, sizeof(float) * n * 32 * 32 * 32));
        hck(hipMalloc(&a2, sizeof(float) * n * 32 * 16 * 16));
        hck(hipMalloc(&a3, sizeof(float) * n * 32 * 8 * 8));
        hck(hipMalloc(&a4, sizeof(float) * n * 32 * 4 * 4));
        hck(hipMalloc(&g1, sizeof(float) * n * 32 * 32 * 32));
        hck(hipMalloc(&g2, sizeof(float) * n * 32 * 16 * 16));
        hck(hipMalloc(&g3, sizeof(float) * n * 32 * 8 * 8));
        hck(hipMalloc(&g4, sizeof(float) * n * 32 * 4 * 4));
        hck(hipMalloc(&y, sizeof(float) * n * 6));
        hck(hipMalloc(&t, sizeof(float) * n * 6));
        hck(hipMalloc(&gy, sizeof(float) * n * 6));
        hck(hipMalloc(&sq, sizeof(float) * n * 6));
        cap = B;
    }
    // the fp16 weight layouts, remade after any parameter write
    void half_weights() {
        if (!wh) hck(hipMalloc(&wh, sizeof(_Float16) * mf::kHalfWeights));
        if (hver == pver) return;
        hipLaunchKernelGGL(k_half_weights, dim3(blocks(mf::kHalfWeights, 256)), dim3(256), 0, st, p, wh);
        hck(hipGetLastError());
        hver = pver;
    }
    void forward(int B) {
        const float* P = p;
        hipLaunchKernelGGL(k_conv_fwd, dim3(blocks((size_t)B * 32 * 32 * 32, 256)), dim3(256), 0, st, x, 2, 64,
                           P + S::off_w(0), P + S::off_b(0), a1, B);
        hipLaunchKernelGGL(k_conv_fwd, dim3(blocks((size_t)B * 32 * 16 * 16, 256)), dim3(256), 0, st, a1, 32, 32,
                           P + S::off_w(1), P + S::off_b(1), a2, B);
        hipLaunchKernelGGL(k_conv_fwd, dim3(blocks((size_t)B * 32 * 8 * 8, 256)), dim3(256), 0, st, a2, 32, 16,
                           P + S::off_w(2), P + S::off_b(2), a3, B);
        hipLaunchKernelGGL(k_conv_fwd, dim3(blocks((size_t)B * 32 * 4 * 4, 256)), dim3(256), 0, st, a3, 32, 8,
                           P + S::off_w(3), P + S::off_b(3), a4, B);
        hipLaunchKernelGGL(k_fc_fwd, dim3(blocks((size_t)B * 6, 256)), dim3(256), 0, st, a4, P + S::off_fw(),
                           P + S::off_fb(), y, B);
        hck(hipGetLastError());
    }
    void backward(int B) {
        const float* P = p;
        hipLaunchKernelGGL(k_loss, dim3(blocks((size_t)B * 6, 256)), dim3(256), 0, st, y, t, gy, sq, B);
        hipLaunchKernelGGL(k_fc_bwd_w, dim3(blocks(6 * 512 + 6, 256)), dim3(256), 0, st, a4, gy, g + S::off_fw(),
                           g + S::off_fb(), B);
        hipLaunchKernelGGL(k_fc_bwd_in, dim3(blocks((size_t)B * 512, 256)), dim3(256), 0, st, gy, P + S::off_fw(),
                           a4, g4, B);
        // layer 3: in a3 (32 x 8 x 8), out g4 (32 x 4 x 4)
        hipLaunchKernelGGL(k_conv_bwd_w, dim3(S::conv_w(3) + 32), dim3(256), 0, st, a3, 32, 8, g4, g + S::off_w(3),
                           g + S::off_b(3), B);
        hipLaunchKernelGGL(k_conv_bwd_in, dim3(blocks((size_t)B * 32 * 8 * 8, 256)), dim3(256), 0, st, g4, 32, 8,
                           P + S::off_w(3), a3, g3, B);
        hipLaunchKernelGGL(k_conv_bwd_w, dim3(S::conv_w(2) + 32), dim3(256), 0, st, a2, 32, 16, g3, g + S::off_w(2),
                           g + S::off_b(2), B);
        hipLaunchKernelGGL(k_conv_bwd_in, dim3(blocks((size_t)B * 32 * 16 * 16, 256)), dim3(256), 0, st, g3, 32, 16,
                           P + S::off_w(2), a2, g2, B);
        hipLaunchKernelGGL(k_conv_bwd_w, dim3(S::conv_w(1) + 32), dim3(256), 0, st, a1, 32, 32, g2, g + S::off_w(1),
                           g + S::off_b(1), B);
        hipLaunchKernelGGL(k_conv_bwd_in, dim3(blocks((size_t)B * 32 * 32 * 32, 256)), dim3(256), 0, st, g2, 32, 32,
                           P + S::off_w(1), a1, g1, B);
        hipLaunchKernelGGL(k_conv_bwd_w, dim3(S::conv_w(0) + 32), dim3(256), 0, st, x, 2, 64, g1, g + S::off_w(0),
                           g + S::off_b(0), B);
        hck(hipGetLastError());
    }
};

Seeder::Seeder() : d_(new Impl) {}
Seeder::~Seeder() {
    if (d_->st) (void)hipStreamDestroy(d_->st);
    delete d_;
}

static bool ensure_stream(hipStream_t& st) {
    return st || hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
}

int Seeder::init(uint64_t seed) {
    try {
        if (!ensure_stream(d_->st)) return -3;
        d_->params_alloc();
        std::vector<float> h(S::total(), 0.0f);
        // weights and biases uniform in +-1 / sqrt(fan-in) (He-normal weights
        // with zero biases leave the loss at the targets' variance here: the
        // regression never starts)
        uint64_t k = 0;
        auto uni = [&](float b) { return b * (2.0f * rnd(seed, 77, k++) - 1.0f); };
        for (int l = 0; l < 4; ++l) {
            const float b = 1.0f / std::sqrt((float)((l == 0 ? S::kInCh : S::kCh) * 9));
            for (int i = 0; i < S::conv_w(l); ++i) h[S::off_w(l) + i] = uni(b);
            for (int i = 0; i < S::kCh; ++i) h[S::off_b(l) + i] = uni(b);
        }
        const float bf = 1.0f / std::sqrt((float)S::kFeat);
        for (int i = 0; i < S::kOut * S::kFeat; ++i) h[S::off_fw() + i] = uni(bf);
        for (int i = 0; i < S::kOut; ++i) h[S::off_fb() + i] = uni(bf);
        hck(hipMemcpyAsync(d_->p, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice, d_->st));
        hck(hipStreamSynchronize(d_->st));
        d_->pver++;
        ready_ = true;
        return 0;
    } catch (const HipFailS&) {
        return -3;
    }
}

namespace {
constexpr char kMagic[8] = {'S', 'D', 'I', 'C', 'S', 'E', 'D', '1'};
}

int Seeder::load(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return -3;
    char mg[8];
    int32_t n = 0;
    f.read(mg, 8);
    f.read(reinterpret_cast<char*>(&n), 4);
    if (!f || std::memcmp(mg, kMagic, 8) != 0 || n != S::total()) return -3;
    std::vector<float> h(n);
    f.read(reinterpret_cast<char*>(h.data()), sizeof(float) * n);
    if (!f) return -3;
    for (float v : h)
        if (!std::isfinite(v)) return -3;
    try {
        if (!ensure_stream(d_->st)) return -3;
        d_->params_alloc();
        hck(hipMemcpyAsync(d_->p, h.data(), sizeof(float) * n, hipMemcpyHostToDevice, d_->st));
        hck(hipStreamSynchronize(d_->st));
        d_->pver++;
    } catch (const HipFailS&) {
        return -3;
    }
    ready_ = true;
    return 0;
}

std::vector<float> Seeder::params() const {
    std::vector<float> h;
    if (!ready_) return h;
    h.resize(S::total());
    if (hipMemcpyAsync(h.data(), d_->p, sizeof(float) * h.size(), hipMemcpyDeviceToHost, d_->st) != hipSuccess ||
        hipStreamSynchronize(d_->st) != hipSuccess)
        h.clear();
    return h;
}

int Seeder::save(const std::string& path) const {
    const std::vector<float> h = params();
    if (h.empty()) return -3;
    std::ofstream f(path, std::ios::binary);
    if (!f) return -3;
    const int32_t n = (int32_t)h.size();
    f.write(kMagic, 8);
    f.write(reinterpret_cast<const char*>(&n), 4);
    f.write(reinterpret_cast<const char*>(h.data()), sizeof(float) * h.size());
    return f ? 0 : -3;
}

int Seeder::train(const uint8_t* ref, int w, int h, const SeederTrainConfig& cfg, std::vector<float>* losses) {
    if (!ready_ || !ref || w < 1024 || h < 1024 || cfg.batch < 1 || cfg.batch > 4096 || cfg.steps < 1) return -3;
    try {
        Impl& d = *d_;
        d.batch_alloc(cfg.batch);
        const int B = cfg.batch;
        std::vector<float> sq(B * 6);
        double acc = 0.0;
        int cnt = 0;
        const float b1 = 0.9f, b2 = 0.999f;
        double p1 = 1.0, p2 = 1.0;
        for (int s = 0; s < cfg.steps; ++s) {
            hipLaunchKernelGGL(k_pairs, dim3(blocks((size_t)B * 4096, 256)), dim3(256), 0, d.st, ref, w, h, B, cfg.seed,
                               (uint64_t)s, cfg.max_shift, cfg.max_grad, cfg.noise,
                               PairBumps{cfg.het_frac, cfg.bump_amp, cfg.bump_min, cfg.bump_max}, d.x, d.t);
            hipLaunchKernelGGL(k_normalize, dim3(2 * B), dim3(256), 0, d.st, d.x);
            d.forward(B);
            d.backward(B);
            p1 *= b1;
            p2 *= b2;
            // cosine decay from lr to lr_end
            const double c = 0.5 * (1.0 + std::cos(3.14159265358979 * s / (double)cfg.steps));
            const float lr = (float)(cfg.lr_end + (cfg.lr - cfg.lr_end) * c);
            d.pver++;
            hipLaunchKernelGGL(k_adam, dim3(blocks(S::total(), 256)), dim3(256), 0, d.st, d.p, d.g, d.m, d.v, S::total(),
                               lr, b1, b2, 1e-8f, (float)(1.0 - p1), (float)(1.0 - p2));
            hck(hipGetLastError());
            if (losses && (s % 20 == 19 || s == cfg.steps - 1)) {
                hck(hipMemcpyAsync(sq.data(), d.sq, sizeof(float) * sq.size(), hipMemcpyDeviceToHost, d.st));
                hck(hipStreamSynchronize(d.st));
                double l = 0.0;
                for (float q : sq) l += q;
                acc += l / (double)sq.size();
                ++cnt;
                if ((s + 1) % 100 == 0 || s == cfg.steps - 1) {
                    losses->push_back((float)(acc / cnt));
                    acc = 0.0;
                    cnt = 0;
                }
            }
        }
        hck(hipStreamSynchronize(d.st));
        return 0;
    } catch (const HipFailS&) {
        return -3;
    }
}

int Seeder::grad(const float* params, const float* x, const float* t, int batch, float* grads, float* loss) {
    if (batch < 1 || batch > 4096) return -3;
    try {
        if (!ensure_stream(d_->st)) return -3;
        Impl& d = *d_;
        d.params_alloc();
        d.batch_alloc(batch);
        hck(hipMemcpyAsync(d.p, params, sizeof(float) * S::total(), hipMemcpyHostToDevice, d.st));
        d.pver++;
        hck(hipMemcpyAsync(d.x, x, sizeof(float) * batch * 2 * 64 * 64, hipMemcpyHostToDevice, d.st));
        hck(hipMemcpyAsync(d.t, t, sizeof(float) * batch * 6, hipMemcpyHostToDevice, d.st));
        d.forward(batch);
        d.backward(batch);
        std::vector<float> sq(batch * 6);
        hck(hipMemcpyAsync(grads, d.g, sizeof(float) * S::total(), hipMemcpyDeviceToHost, d.st));
        hck(hipMemcpyAsync(sq.data(), d.sq, sizeof(float) * sq.size(), hipMemcpyDeviceToHost, d.st));
        hck(hipStreamSynchronize(d.st));
        double l = 0.0;
        for (float q : sq) l += q;
        *loss = (float)(l / (double)sq.size());
        ready_ = true;
        return 0;
    } catch (const HipFailS&) {
        return -3;
    }
}

int Seeder::forward(const float* params, const float* x, int batch, float* y) {
    if (batch < 1 || batch > 4096) return -3;
    try {
        if (!ensure_stream(d_->st)) return -3;
        Impl& d = *d_;
        d.params_alloc();
        d.batch_alloc(batch);
        hck(hipMemcpyAsync(d.p, params, sizeof(float) * S::total(), hipMemcpyHostToDevice, d.st));
        d.pver++;
        hck(hipMemcpyAsync(d.x, x, sizeof(float) * batch * 2 * 64 * 64, hipMemcpyHostToDevice, d.st));
        d.forward(batch);
        hck(hipMemcpyAsync(y, d.y, sizeof(float) * batch * 6, hipMemcpyDeviceToHost, d.st));
        hck(hipStreamSynchronize(d.st));
        ready_ = true;
        return 0;
    } catch (const HipFailS&) {
        return -3;
    }
}

int Seeder::forward_mfma(const float* params, const float* x, int batch, float* y) {
    if (batch < 1 || batch > 4096) return -3;
    try {
        if (!ensure_stream(d_->st)) return -3;
        Impl& d = *d_;
        d.params_alloc();
        d.batch_alloc(batch);
        hck(hipMemcpyAsync(d.p, params, sizeof(float) * S::total(), hipMemcpyHostToDevice, d.st));
        d.pver++;
        hck(hipMemcpyAsync(d.x, x, sizeof(float) * batch * 2 * 64 * 64, hipMemcpyHostToDevice, d.st));
        d.half_weights();
        hipLaunchKernelGGL(k_seed_mfma<false>, dim3(batch), dim3(512), 0, d.st, nullptr, nullptr, 0, 0, nullptr,
                           nullptr, d.x, d.wh, d.p, d.y);
        hck(hipGetLastError());
        hck(hipMemcpyAsync(y, d.y, sizeof(float) * batch * 6, hipMemcpyDeviceToHost, d.st));
        hck(hipStreamSynchronize(d.st));
        ready_ = true;
        return 0;
    } catch (const HipFailS&) {
        return -3;
    }
}

int Seeder::infer(const float* r2, const float* d2, int w2, int h2, const int* c, const int* off, int n, float* out,
                  hipStream_t st) {
    if (!ready_ || n <= 0) return n == 0 ? 0 : -3;
    try {
        Impl& d = *d_;
        d.batch_alloc(n);
        if (n > d.icap) {
            if (d.ci) (void)hipFree(d.ci);
            if (d.oi) (void)hipFree(d.oi);
            d.ci = d.oi = nullptr;
            hck(hipMalloc(&d.ci, sizeof(int) * 2 * n));
            hck(hipMalloc(&d.oi, sizeof(int) * 2 * n));
            d.icap = n;
        }
        // on the caller's stream: the pyramids were made there
        struct StreamSwap {
            hipStream_t& s;
            hipStream_t keep;
            StreamSwap(hipStream_t& x, hipStream_t to) : s(x), keep(x) { s = to; }
            ~StreamSwap() { s = keep; }
        } swap(d.st, st);
        hck(hipMemcpyAsync(d.ci, c, sizeof(int) * 2 * n, hipMemcpyHostToDevice, st));
        hck(hipMemcpyAsync(d.oi, off, sizeof(int) * 2 * n, hipMemcpyHostToDevice, st));
        if (use_mfma) {
            d.half_weights();
            hipLaunchKernelGGL(k_seed_mfma<true>, dim3(n), dim3(512), 0, st, r2, d2, w2, h2, d.ci, d.oi, nullptr, d.wh,
                               d.p, d.y);
            hck(hipGetLastError());
        } else {
            hipLaunchKernelGGL(k_cut, dim3(blocks((size_t)n * 4096, 256)), dim3(256), 0, st, r2, d2, w2, h2, d.ci,
                               d.oi, n, d.x);
            hipLaunchKernelGGL(k_normalize, dim3(2 * n), dim3(256), 0, st, d.x);
            d.forward(n);
        }
        std::vector<float> y((size_t)n * 6);
        hck(hipMemcpyAsync(y.data(), d.y, sizeof(float) * y.size(), hipMemcpyDeviceToHost, st));
        hck(hipStreamSynchronize(st));
        for (int k = 0; k < n; ++k) {
            out[6 * k + 0] = y[6 * k + 0] * S::kUvScale;
            out[6 * k + 1] = y[6 * k + 1] * S::kUvScale;
            for (int q = 2; q < 6; ++q) out[6 * k + q] = y[6 * k + q] * S::kGradScale;
        }
        return 0;
    } catch (const HipFailS&) {
        return -3;
    }
}

}  // namespace sdic
