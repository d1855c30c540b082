// Reference precompute kernels for gfx950 (HBM-bound):
//   k_prep_ref   u8 -> float intensity plane + 4th-order gradients, Ghost Wall
//                (image_processor.cpp:9-64, reference_cache.cpp:18-35)
//   k_scan       grid scan: boundary buffer and the strict per-pixel mask test
//                (full_field_solver.cpp:271-317), bit-exact integer work
//   k_pool       Hessian pool, one wave per grid node
//                (subset_precomputer.cpp:42-149, hessian.hpp:40-52)
//   k_down*      2x2 box pyramid of the global-motion search
//   k_patch      coarse-to-fine ZNCC patch search of the global-motion model
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "solver.hpp"
#include "wave.hpp"

#pragma clang fp contract(off)

namespace sdic {

namespace {

// (-I[+2] + 8 I[+1] - 8 I[-1] + I[-2]) / 12 with the reference's evaluation
// order; zero on the 2 px border.
template <typename T>
__device__ __forceinline__ void grad_at(const T* img, int w, int h, int x, int y, float& gx, float& gy) {
    gx = 0.0f;
    gy = 0.0f;
    if (x < 2 || y < 2 || x >= w - 2 || y >= h - 2) return;
    const int i = y * w + x;
    gx = (-(float)img[i + 2] + 8.0f * (float)img[i + 1] - 8.0f * (float)img[i - 1] + (float)img[i - 2]) / 12.0f;
    gy = (-(float)img[i + 2 * w] + 8.0f * (float)img[i + w] - 8.0f * (float)img[i - w] +
          (float)img[i - 2 * w]) /
         12.0f;
}

template <typename T>
__global__ void k_prep_ref(const T* __restrict__ img, const uint8_t* __restrict__ mask, int w, int h,
                           float* __restrict__ I, float* __restrict__ gx, float* __restrict__ gy) {
    const int64_t n = (int64_t)w * h;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % w), y = (int)(i / w);
        float a, b;
        grad_at(img, w, h, x, y, a, b);
        gx[i] = a;
        gy[i] = b;
        // The Ghost Wall is stamped after the gradients, which see the real pixels.
        I[i] = (mask && mask[i] < 128) ? kGhostWallValue : (float)img[i];
    }
}

// One thread per grid node (no mask): skip = inside the boundary buffer
// d/2 + 4 + 15. The kept nodes are counted by a ballot per wave and a sum
// per 1024-node workgroup: one same-address atomic per workgroup.
constexpr int kScanBlock = 1024;
__global__ __launch_bounds__(kScanBlock) void k_scan(int w, int h, int rect_x, int rect_y, int step, int gw, int gh,
                                                     int dim, const int* __restrict__ dims,
                                                     uint8_t* __restrict__ skip, int* count) {
    __shared__ int kept_w[kScanBlock / 64];
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = node < gw * gh;
    bool s = true;
    if (live) {
        const int rx = rect_x + (node % gw) * step, ry = rect_y + (node / gw) * step;
        const int half = (dims ? dims[node] : dim) / 2, buf = half + 4 + 15;
        s = rx - buf < 0 || rx + buf >= w || ry - buf < 0 || ry + buf >= h;
        skip[node] = s ? 1 : 0;
    }
    const uint64_t kept = __ballot(live && !s);
    if ((threadIdx.x & 63) == 0) kept_w[threadIdx.x >> 6] = (int)__popcll(kept);
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int k = 0; k < kScanBlock / 64; ++k) t += kept_w[k];
        if (t) atomicAdd(count, t);
    }
}

// The same with a mask: skip also when any subset pixel is masked (< 128).
// A workgroup takes 64 consecutive nodes, each of its four waves every
// fourth of them, and a node's d x d window is read by the wave's 64 lanes
// together; one count per workgroup.
constexpr int kScanNodesPerGroup = 64;
__global__ __launch_bounds__(256) void k_scan_masked(const uint8_t* __restrict__ mask, int w, int h, int rect_x,
                                                     int rect_y, int step, int gw, int gh, int dim,
                                                     const int* __restrict__ dims, uint8_t* __restrict__ skip,
                                                     int* count) {
    __shared__ int kept_w[4];
    const int lane = (int)(threadIdx.x & 63), wid = (int)(threadIdx.x >> 6);
    const int n = gw * gh, base = (int)blockIdx.x * kScanNodesPerGroup;
    int kept = 0;
    for (int k = wid; k < kScanNodesPerGroup; k += 4) {
        const int node = base + k;
        if (node >= n) break;
        const int rx = rect_x + (node % gw) * step, ry = rect_y + (node / gw) * step;
        const int d = dims ? dims[node] : dim, half = d / 2, buf = half + 4 + 15;
        bool s = rx - buf < 0 || rx + buf >= w || ry - buf < 0 || ry + buf >= h;
        if (!s) {
            // pixel i of the window at (q, r) = divmod(i, d), advanced by divmod(64, d)
            const int q64 = 64 / d, r64 = 64 % d;
            int q = lane / d, r = lane % d;
            bool bad = false;
            for (int i = lane; i < d * d; i += 64) {
                bad |= mask[(int64_t)(ry + q - half) * w + rx + r - half] < 128;
                r += r64;
                q += q64;
                if (r >= d) {
                    r -= d;
                    q++;
                }
            }
            s = __ballot(bad) != 0;
        }
        if (lane == 0) skip[node] = s ? 1 : 0;
        kept += s ? 0 : 1;
    }
    if (lane == 0) kept_w[wid] = kept;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int t = kept_w[0] + kept_w[1] + kept_w[2] + kept_w[3];
        if (t) atomicAdd(count, t);
    }
}

// The reference planes a pool pass reads: float gradients, or their fp16
// copies (bits in uint16_t) when the gradient storage is fp16.
struct GradRead {
    const float* gx;
    const float* gy;
    const uint16_t* gx16;
    const uint16_t* gy16;
    bool bf16;   // the 16-bit copies are bf16 (else fp16)
    __device__ __forceinline__ float x(int a) const { return gx16 ? half_bits_to_float(gx16[a], bf16) : gx[a]; }
    __device__ __forceinline__ float y(int a) const { return gy16 ? half_bits_to_float(gy16[a], bf16) : gy[a]; }
};

typedef _Float16 sdic_half4 __attribute__((ext_vector_type(4)));
typedef float sdic_float4 __attribute__((ext_vector_type(4)));

// compute_hessian_only's pixel passes for the subset at (cx, cy): the mean
// and the standard deviation over the valid pixels, then the 21 Hessian sums,
// either as FP32 lane sums with wave reductions or (MFMA) as one 16x16x16 fp16
// MFMA per 16 pixels with FP32 accumulation. False when the subset is invalid
// (outside the image, or under half of it valid). Results are wave-uniform.
template <bool MFMA>
__device__ bool pool_sums(const float* __restrict__ I, GradRead G, int w, int h, int cx, int cy, int dim,
                          float* lds36, float* hh, float& mean, float& inv) {
    const int lane = threadIdx.x & 63;
    const int half = dim / 2, n = dim * dim;
    if (dim < 3 || (dim & 1) == 0 || cx - half < 0 || cx + half >= w || cy - half < 0 || cy + half >= h)
        return false;
    // pixel i = lane + 64 k of the subset at (q, r) = divmod(i, dim), advanced
    // by divmod(64, dim) per step instead of divided anew; each pass reads
    // kChunk of a lane's pixels before it uses any (their loads in flight
    // together) and adds them in pixel order
    constexpr int kChunk = 4;
    const int q0 = lane / dim, r0 = lane % dim, q64 = 64 / dim, r64 = 64 % dim;
    const int base = (cy - half) * w + cx - half;
    const auto offsets = [&](int i, int& q, int& r, int (&off)[kChunk], int (&oxy)[kChunk][2], bool (&in)[kChunk]) {
#pragma unroll
        for (int c = 0; c < kChunk; ++c) {
            in[c] = i + 64 * c < n;
            off[c] = in[c] ? q * w + r : 0;
            oxy[c][0] = r - half;
            oxy[c][1] = q - half;
            r += r64;
            q += q64;
            if (r >= dim) {
                r -= dim;
                q++;
            }
        }
    };
    if constexpr (!MFMA) {
        // a subset of up to 64 kHeld pixels: its intensities and gradients
        // loaded into registers at once, then the three passes over them
        constexpr int kHeld = 16;
        if (n <= 64 * kHeld) {
            float v[kHeld], gxv[kHeld], gyv[kHeld];
            int oxy[kHeld][2];
            bool in[kHeld];
            {
                int q = q0, r = r0;
#pragma unroll
                for (int c = 0; c < kHeld; ++c) {
                    in[c] = lane + 64 * c < n;
                    const int off = in[c] ? q * w + r : 0;
                    oxy[c][0] = r - half;
                    oxy[c][1] = q - half;
                    v[c] = I[base + off];
                    gxv[c] = G.x(base + off);
                    gyv[c] = G.y(base + off);
                    r += r64;
                    q += q64;
                    if (r >= dim) {
                        r -= dim;
                        q++;
                    }
                }
            }
            float sum = 0.0f;
            int cnt = 0;
#pragma unroll
            for (int c = 0; c < kHeld; ++c) {
                if (!in[c] || v[c] < kGhostWallThreshold) continue;
                sum += v[c];
                cnt++;
            }
            sum = wave_sum(sum);
            cnt = wave_sum_i(cnt);
            if (cnt < n * 0.5f) return false;
            mean = sum / (float)cnt;
            float ss = 0.0f;
#pragma unroll
            for (int c = 0; c < kHeld; ++c) {
                if (!in[c] || v[c] < kGhostWallThreshold) continue;
                const float d = v[c] - mean;
                ss += d * d;
            }
            ss = wave_sum(ss);
            float sd = sqrtf(ss / (float)cnt);
            if (sd < 1e-5f) sd = 1.0f;
            inv = 1.0f / sd;
#pragma unroll
            for (int k = 0; k < 21; ++k) hh[k] = 0.0f;
#pragma unroll
            for (int c = 0; c < kHeld; ++c) {
                if (!in[c] || v[c] < kGhostWallThreshold) continue;
                const float g0 = gxv[c] * inv, g1 = gyv[c] * inv;
                const float fx = (float)oxy[c][0], fy = (float)oxy[c][1];
                const float s6[6] = {g0, g1, g0 * fx, g0 * fy, g1 * fx, g1 * fy};
                int k = 0;
#pragma unroll
                for (int rr = 0; rr < 6; ++rr)
#pragma unroll
                    for (int cc = rr; cc < 6; ++cc) hh[k++] += s6[rr] * s6[cc];
            }
#pragma unroll
            for (int k = 0; k < 21; ++k) hh[k] = wave_sum(hh[k]);
            return true;
        }
    }
    float sum = 0.0f;
    int cnt = 0;
    for (int i = lane, q = q0, r = r0; i < n; i += 64 * kChunk) {
        int off[kChunk], oxy[kChunk][2];
        bool in[kChunk];
        offsets(i, q, r, off, oxy, in);
        float v[kChunk];
#pragma unroll
        for (int c = 0; c < kChunk; ++c) v[c] = I[base + off[c]];
#pragma unroll
        for (int c = 0; c < kChunk; ++c) {
            if (!in[c] || v[c] < kGhostWallThreshold) continue;
            sum += v[c];
            cnt++;
        }
    }
    sum = wave_sum(sum);
    cnt = wave_sum_i(cnt);
    if (cnt < n * 0.5f) return false;
    mean = sum / (float)cnt;
    float ss = 0.0f;
    for (int i = lane, q = q0, r = r0; i < n; i += 64 * kChunk) {
        int off[kChunk], oxy[kChunk][2];
        bool in[kChunk];
        offsets(i, q, r, off, oxy, in);
        float v[kChunk];
#pragma unroll
        for (int c = 0; c < kChunk; ++c) v[c] = I[base + off[c]];
#pragma unroll
        for (int c = 0; c < kChunk; ++c) {
            if (!in[c] || v[c] < kGhostWallThreshold) continue;
            const float d = v[c] - mean;
            ss += d * d;
        }
    }
    ss = wave_sum(ss);
    float sd = sqrtf(ss / (float)cnt);
    if (sd < 1e-5f) sd = 1.0f;
    inv = 1.0f / sd;
    if constexpr (!MFMA) {
#pragma unroll
        for (int k = 0; k < 21; ++k) hh[k] = 0.0f;
        for (int i = lane, q = q0, r = r0; i < n; i += 64 * kChunk) {
            int off[kChunk], oxy[kChunk][2];
            bool in[kChunk];
            offsets(i, q, r, off, oxy, in);
            float v[kChunk], gxv[kChunk], gyv[kChunk];
#pragma unroll
            for (int c = 0; c < kChunk; ++c) {
                v[c] = I[base + off[c]];
                gxv[c] = G.x(base + off[c]);
                gyv[c] = G.y(base + off[c]);
            }
#pragma unroll
            for (int c = 0; c < kChunk; ++c) {
                if (!in[c] || v[c] < kGhostWallThreshold) continue;
                const float g0 = gxv[c] * inv, g1 = gyv[c] * inv;
                const float fx = (float)oxy[c][0], fy = (float)oxy[c][1];
                const float s6[6] = {g0, g1, g0 * fx, g0 * fy, g1 * fx, g1 * fy};
                int k = 0;
#pragma unroll
                for (int rr = 0; rr < 6; ++rr)
#pragma unroll
                    for (int cc = rr; cc < 6; ++cc) hh[k++] += s6[rr] * s6[cc];
            }
        }
#pragma unroll
        for (int k = 0; k < 21; ++k) hh[k] = wave_sum(hh[k]);
    } else {
        // A (16 x 16, rows = the six steepest-descent components, columns = 16
        // pixels) times its transpose: lane l supplies component l % 16 of the
        // pixels 4 (l / 16) .. 4 (l / 16) + 3 of the chunk, which is its share
        // of both operands; D[r][c] comes back in lane 16 (r / 4) + c, slot r % 4.
        const int comp = lane & 15, grp = lane >> 4;
        sdic_float4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int c0 = 0; c0 < n; c0 += 16) {
            sdic_half4 av;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int p = c0 + 4 * grp + j;
                float s = 0.0f;
                if (comp < 6 && p < n) {
                    const int oy = p / dim - half, ox = p % dim - half;
                    const int a = (cy + oy) * w + cx + ox;
                    if (!(I[a] < kGhostWallThreshold)) {
                        const float g0 = G.x(a) * inv, g1 = G.y(a) * inv;
                        const float gs = comp == 0 || comp == 2 || comp == 3 ? g0 : g1;
                        const float f = (comp == 2 || comp == 4) ? (float)ox : (float)oy;
                        s = comp < 2 ? gs : gs * f;
                    }
                }
                av[j] = (_Float16)s;
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16f16(av, av, acc, 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = 4 * grp + j;
            if (r < 6 && comp < 6) lds36[r * 6 + comp] = acc[j];
        }
        __syncthreads();
        int k = 0;
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int c = r; c < 6; ++c) hh[k++] = lds36[r * 6 + c];
        __syncthreads();
    }
    return true;
}

// compute_hessian_only's 6x6 part from the 21 sums (subset_precomputer.cpp:
// 104-149): the 2x2 condition test, hessian::well_conditioned, the inverse,
// and under LM the damped matrix's test and inverse (optimization_engine.cpp:
// 302-315). step receives the step matrix; Lp (optional) the FP64 Cholesky
// factor of the step matrix's source, its first entry NaN when it has none.
// Returns valid.
template <bool CHOL>
__device__ __forceinline__ int pool_tail(const float* hh, int half, bool lm, float* step, double* Lp) {
    float H[36];
    {
        int k = 0;
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int c = r; c < 6; ++c) {
                H[r * 6 + c] = hh[k];
                H[c * 6 + r] = hh[k];
                ++k;
            }
    }
    const float det2 = H[0] * H[7] - H[6] * H[1];
    const float nrm2 = H[0] * H[0] + H[1] * H[1] + H[6] * H[6] + H[7] * H[7];
    const float cond2 = (fabsf(det2) > 1e-12f) ? (nrm2 / fabsf(det2)) : 1.0e13f;
    int valid = 1;
    bool damped = false;
    if (cond2 > 1.0e12f || !well_conditioned_reg(H, half)) {
        valid = 0;
    } else {
        gj_inverse_reg(H, step);
        if (lm) {
            float Hd[36];
#pragma unroll
            for (int q = 0; q < 36; ++q) Hd[q] = H[q];
            Hd[0] += kLmAlpha;
            Hd[7] += kLmAlpha;
            if (well_conditioned_reg(Hd, half)) {
                gj_inverse_reg(Hd, step);
                damped = true;
            }
        }
    }
    if constexpr (CHOL) {
#pragma unroll
        for (int q = 0; q < kCholStride; ++q) Lp[q] = 0.0;
        bool ok = false;
        if (valid) {
            if (damped) {
                H[0] += kLmAlpha;
                H[7] += kLmAlpha;
            }
            ok = chol64_factor(H, Lp);
        }
        if (!ok) Lp[0] = __builtin_nan("");
    }
    return valid;
}

// The grid pool in two launches: k_pool_sums, one wave per node, writes the
// sums and statistics; k_pool_tail, one thread per node, does the serial 6x6
// part, so a wave finishes 64 nodes' algebra at once.
struct alignas(16) PoolRaw {
    float hh[21];
    float mean, inv;
    int32_t ok;
};

template <bool MFMA>
__global__ __launch_bounds__(64) void k_pool_sums(const float* __restrict__ I, GradRead G, int w, int h,
                                                  const uint8_t* __restrict__ skip, int gw, int gh, int rect_x,
                                                  int rect_y, int step, int dim, const int* __restrict__ dims,
                                                  PoolRaw* __restrict__ raw) {
    __shared__ float lds36[36];
    const int node = xcd_remap(blockIdx.x, gridDim.x);
    if (node >= gw * gh) return;
    const int lane = threadIdx.x & 63;
    float hh[21], mean = 0.0f, inv = 1.0f;
    const bool ok = !skip[node] && pool_sums<MFMA>(I, G, w, h, rect_x + (node % gw) * step,
                                                   rect_y + (node / gw) * step, dims ? dims[node] : dim, lds36, hh,
                                                   mean, inv);
    float v = 0.0f;
    if (ok) {
#pragma unroll
        for (int k = 0; k < 21; ++k) v = lane == k ? hh[k] : v;
        v = lane == 21 ? mean : v;
        v = lane == 22 ? inv : v;
    }
    float* dst = reinterpret_cast<float*>(raw + node);
    if (lane < 23) dst[lane] = v;
    if (lane == 23) reinterpret_cast<int32_t*>(dst)[23] = ok ? 1 : 0;
}

template <bool CHOL>
__global__ __launch_bounds__(256) void k_pool_tail(const PoolRaw* __restrict__ raw, int n, int dim,
                                                   const int* __restrict__ dims, int lm, PoolEntry* __restrict__ pool,
                                                   double* __restrict__ chol) {
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= n) return;
    const int half = (dims ? dims[node] : dim) / 2;
    PoolEntry* dst = pool + node;
    if (!raw[node].ok) {
        dst->valid = 0;
        if constexpr (CHOL) chol[(size_t)node * kCholStride] = __builtin_nan("");
        return;
    }
    float hh[21];
#pragma unroll
    for (int k = 0; k < 21; ++k) hh[k] = raw[node].hh[k];
    float step[36];
    double Lp[kCholStride];
    const int valid = pool_tail<CHOL>(hh, half, lm != 0, step, Lp);
    if (valid) {
#pragma unroll
        for (int q = 0; q < 36; ++q) dst->hstep[q] = step[q];
    }
    dst->mean = raw[node].mean;
    dst->inv_std = raw[node].inv;
    dst->valid = valid;
    if constexpr (CHOL) {
#pragma unroll
        for (int q = 0; q < kCholStride; ++q) chol[(size_t)node * kCholStride + q] = Lp[q];
    }
}

// One launch per node list (anchor canvases, the node solver's pool): the
// sums and the 6x6 part in one wave, the latter evaluated redundantly by every
// lane, lane q writing entry q.
__global__ __launch_bounds__(64) void k_pool_list(const float* __restrict__ I, GradRead G, int w, int h,
                                                  const int* __restrict__ cxy, int n, int dim, int lm,
                                                  PoolEntry* pool) {
    __shared__ float lds36[36];
    const int k = blockIdx.x;
    if (k >= n) return;
    const int lane = threadIdx.x & 63;
    PoolEntry* dst = pool + k;
    float hh[21], mean = 0.0f, inv = 1.0f;
    if (!pool_sums<false>(I, G, w, h, cxy[2 * k], cxy[2 * k + 1], dim, lds36, hh, mean, inv)) {
        if (lane == 0) dst->valid = 0;
        return;
    }
    float step[36];
    const int valid = pool_tail<false>(hh, dim / 2, lm != 0, step, nullptr);
    if (valid) {
        float v = 0.0f;
#pragma unroll
        for (int q = 0; q < 36; ++q) v = lane == q ? step[q] : v;
        if (lane < 36) dst->hstep[lane] = v;
    }
    if (lane == 0) {
        dst->mean = mean;
        dst->inv_std = inv;
        dst->valid = valid;
    }
}

template <typename T>
__global__ void k_down(const T* __restrict__ src, int w, int h, float* __restrict__ dst) {
    const int nw = w / 2, nh = h / 2;
    const int64_t n = (int64_t)nw * nh;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % nw), y = (int)(i / nw);
        const T* a = src + (int64_t)(2 * y) * w + 2 * x;
        dst[i] = 0.25f * (((float)a[0] + (float)a[1]) + ((float)a[w] + (float)a[w + 1]));
    }
}

// ---- global-motion patch search (the seeding front end) ---------------------

constexpr int kPatchHalf = 16;
constexpr int kFineRadius = 2;

__device__ __forceinline__ float lv_at(const PyrDev& p, bool def, int l, int idx) {
    if (l == 0) return (float)(def ? p.d0[idx] : p.r0[idx]);
    return def ? p.d[l][idx] : p.r[l][idx];
}

// ZNCC of the 32 x 32 template at (cx, cy) on level l against the window
// shifted by (sx, sy), evaluated by one wave (lanes stride over the template);
// false if either leaves the level's image. Every pyramid value is a dyadic
// rational with at most 6 fractional bits and at most 8 integer bits, so the
// five sums are exact in double in any order: the score does not depend on
// how the lanes split the work.
// The deformed level bilinearly at (x, y), inside [0, w - 1) x [0, h - 1).
__device__ __forceinline__ double lv_bilinear(const PyrDev& p, int l, double x, double y) {
    const int w = p.w[l];
    const int xi = (int)floor(x), yi = (int)floor(y);
    const double fx = x - xi, fy = y - yi;
    const int i = yi * w + xi;
    const double a = lv_at(p, true, l, i), b = lv_at(p, true, l, i + 1);
    const double c = lv_at(p, true, l, i + w), d = lv_at(p, true, l, i + w + 1);
    return (a * (1.0 - fx) + b * fx) * (1.0 - fy) + (c * (1.0 - fx) + d * fx) * fy;
}

__device__ bool zncc(const PyrDev& p, int l, int cx, int cy, int sx, int sy, double& out, double* tstd) {
    const int w = p.w[l], h = p.h[l];
    const int x0 = cx - kPatchHalf, y0 = cy - kPatchHalf, x1 = cx + kPatchHalf, y1 = cy + kPatchHalf;
    if (x0 < 0 || y0 < 0 || x1 > w || y1 > h) return false;
    if (!p.warp) {
        if (x0 + sx < 0 || y0 + sy < 0 || x1 + sx > w || y1 + sy > h) return false;
    } else {
        // the warped template's corners, with the bilinear footprint, inside
        for (int k = 0; k < 4; ++k) {
            const double ox = (k & 1) ? kPatchHalf - 1 : -kPatchHalf, oy = (k & 2) ? kPatchHalf - 1 : -kPatchHalf;
            const double qx = cx + sx + p.f[0] * ox + p.f[1] * oy, qy = cy + sy + p.f[2] * ox + p.f[3] * oy;
            if (!(qx >= 0.0 && qy >= 0.0 && qx < (double)(w - 1) && qy < (double)(h - 1))) return false;
        }
    }
    const int lane = threadIdx.x & 63, side = 2 * kPatchHalf;
    double sr = 0, sd = 0, srr = 0, sdd = 0, srd = 0;
    for (int k = lane; k < side * side; k += 64) {
        const int y = y0 + k / side, x = x0 + k % side;
        const double a = lv_at(p, false, l, y * w + x);
        double b;
        if (!p.warp) {
            b = lv_at(p, true, l, (y + sy) * w + x + sx);
        } else {
            const double ox = x - cx, oy = y - cy;
            b = lv_bilinear(p, l, cx + sx + p.f[0] * ox + p.f[1] * oy, cy + sy + p.f[2] * ox + p.f[3] * oy);
        }
        sr += a;
        sd += b;
        srr += a * a;
        sdd += b * b;
        srd += a * b;
    }
    sr = wave_sum_d(sr);
    sd = wave_sum_d(sd);
    srr = wave_sum_d(srr);
    sdd = wave_sum_d(sdd);
    srd = wave_sum_d(srd);
    const double n = (double)(side * side);
    const double vr = srr - sr * sr / n, vd = sdd - sd * sd / n;
    if (tstd) *tstd = sqrt(fmax(0.0, vr / n));
    if (!(vr > 1e-9) || !(vd > 1e-9)) {
        out = -2.0;
        return true;
    }
    out = (srd - sr * sd / n) / (sqrt(vr) * sqrt(vd));
    return true;
}

struct Cand {
    double s;
    long r2;
    int order;
    int sx, sy;
    int any;
};

// better(a, b): higher score, then smaller |shift|^2, then earlier scan order
__device__ __forceinline__ bool better(const Cand& a, const Cand& b) {
    if (!a.any) return false;
    if (!b.any) return true;
    if (a.s != b.s) return a.s > b.s;
    if (a.r2 != b.r2) return a.r2 < b.r2;
    return a.order < b.order;
}

// One block of 16 waves per patch. Each level's candidates are spread over the
// waves (one candidate per wave at a time); the best by better() is reduced
// over the waves, and better() is a total order, so the choice is the serial
// scan's.
constexpr int kPatchWaves = 16;
// The coarsest level holds most of the candidates ((2 r + 1)^2 against 25 at
// each finer level): it is shared by kPatchSplit blocks per patch, so that
// the GPU's CUs, not one per patch, scan it. Each block keeps the best of its
// share by better(), and the patch's block takes the best of those, the same
// candidate as one block's scan (a total order has one maximum).
constexpr int kPatchSplit = 4;

__global__ __launch_bounds__(64 * kPatchWaves) void k_patch_coarse(PyrDev p, const int* __restrict__ centres,
                                                                   Cand* __restrict__ cand) {
    __shared__ Cand red[kPatchWaves];
    const int pi = blockIdx.x, part = blockIdx.y, lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int px = centres[2 * pi], py = centres[2 * pi + 1];
    const int l = p.levels;
    const int rad = p.coarse_radius, side = 2 * rad + 1, ns = side * side;
    const int csx = p.start ? p.start[2 * pi] : 0, csy = p.start ? p.start[2 * pi + 1] : 0;
    Cand best;
    best.any = 0;
    for (int k = part * kPatchWaves + wid; k < ns; k += kPatchWaves * kPatchSplit) {
        const int sy = csy - rad + k / side, sx = csx - rad + k % side;
        double sc;
        if (!zncc(p, l, px >> l, py >> l, sx, sy, sc, nullptr)) continue;
        Cand c;
        c.s = sc;
        c.r2 = (long)sx * sx + (long)sy * sy;
        c.order = k;
        c.sx = sx;
        c.sy = sy;
        c.any = 1;
        if (better(c, best)) best = c;
    }
    if (lane == 0) red[wid] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        Cand b = red[0];
        for (int k = 1; k < kPatchWaves; ++k)
            if (better(red[k], b)) b = red[k];
        cand[pi * kPatchSplit + part] = b;
    }
}

// coarse: the coarsest level's candidates from k_patch_coarse, or null (this
// block scans that level too)
__global__ __launch_bounds__(64 * kPatchWaves) void k_patch(PyrDev p, const int* __restrict__ centres,
                                                            const Cand* __restrict__ coarse, PatchOut* out) {
    __shared__ Cand red[kPatchWaves];
    __shared__ int s_sx, s_sy, s_ok;
    __shared__ double s_best;
    const int pi = blockIdx.x, lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int px = centres[2 * pi], py = centres[2 * pi + 1];
    const int L = p.levels;
    if (threadIdx.x == 0) {
        s_sx = 0;
        s_sy = 0;
        s_ok = 1;
        s_best = -3.0;
    }
    __syncthreads();
    if (threadIdx.x == 0 && p.start) {
        s_sx = p.start[2 * pi];
        s_sy = p.start[2 * pi + 1];
    }
    __syncthreads();
    for (int l = L; l >= 0; --l) {
        if (!s_ok) break;
        if (l == L && coarse) {
            if (threadIdx.x == 0) {
                Cand b = coarse[pi * kPatchSplit];
                for (int k = 1; k < kPatchSplit; ++k)
                    if (better(coarse[pi * kPatchSplit + k], b)) b = coarse[pi * kPatchSplit + k];
                if (!b.any) {
                    s_ok = 0;
                } else {
                    s_best = b.s;
                    s_sx = b.sx * (l > 0 ? 2 : 1);
                    s_sy = b.sy * (l > 0 ? 2 : 1);
                }
            }
            __syncthreads();
            continue;
        }
        const int rad = (l == L) ? p.coarse_radius : kFineRadius;
        const int side = 2 * rad + 1, ns = side * side;
        const int csx = s_sx, csy = s_sy;
        Cand best;
        best.any = 0;
        for (int k = wid; k < ns; k += kPatchWaves) {
            const int sy = csy - rad + k / side, sx = csx - rad + k % side;
            double sc;
            if (!zncc(p, l, px >> l, py >> l, sx, sy, sc, nullptr)) continue;
            Cand c;
            c.s = sc;
            c.r2 = (long)sx * sx + (long)sy * sy;
            c.order = k;
            c.sx = sx;
            c.sy = sy;
            c.any = 1;
            if (better(c, best)) best = c;
        }
        if (lane == 0) red[wid] = best;
        __syncthreads();
        if (threadIdx.x == 0) {
            Cand b = red[0];
            for (int k = 1; k < kPatchWaves; ++k)
                if (better(red[k], b)) b = red[k];
            if (!b.any) {
                s_ok = 0;
            } else {
                s_best = b.s;
                s_sx = b.sx * (l > 0 ? 2 : 1);
                s_sy = b.sy * (l > 0 ? 2 : 1);
            }
        }
        __syncthreads();
    }
    // level-0 statistics and the four neighbours of the best shift, one wave each
    __shared__ double s_vals[5];
    __shared__ int s_has[5];
    if (s_ok && wid < 5) {
        const int dx[5] = {0, -1, 1, 0, 0}, dy[5] = {0, 0, 0, -1, 1};
        double v = 0.0, t = 0.0;
        const bool ok = zncc(p, 0, px, py, s_sx + dx[wid], s_sy + dy[wid], v, wid == 0 ? &t : nullptr);
        if (lane == 0) {
            s_vals[wid] = wid == 0 ? t : v;
            s_has[wid] = ok ? 1 : 0;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        PatchOut o = {};
        o.ok = s_ok;
        o.sx = s_sx;
        o.sy = s_sy;
        o.best = s_best;
        if (s_ok) {
            o.tstd = s_vals[0];
            o.xm = s_vals[1];
            o.xp = s_vals[2];
            o.ym = s_vals[3];
            o.yp = s_vals[4];
            o.has_x = s_has[1] && s_has[2];
            o.has_y = s_has[3] && s_has[4];
        }
        out[pi] = o;
    }
}

// ---- rotation-invariant patch matching ---------------------------------------
//
// A disc of radius kCanonR at level l is described independently of its
// rotation: its orientation is the angle of its intensity centroid (the
// first moments of the disc about its centre; a disc turned by t has its
// centroid turned by t), and the disc is sampled bilinearly on a polar grid
// (radii 2 .. kCanonR - 1, kCanonAngles angles) turned by that angle, then
// normalised to zero mean and unit norm. Two discs that differ by a rotation
// give nearly the same samples; the score of a pair is their dot product.

__device__ __forceinline__ float lvf(const PyrDev& p, bool def, int l, int idx) {
    return (float)lv_at(p, def, l, idx);
}

// orientation of the disc at integer (x, y): its centroid's angle
__device__ float canon_theta(const PyrDev& p, bool def, int l, int x, int y) {
    const int w = p.w[l];
    float mx = 0.0f, my = 0.0f;
    for (int dy = -kCanonR; dy <= kCanonR; ++dy) {
        const int span = (int)floorf(sqrtf((float)(kCanonR * kCanonR - dy * dy)));
        const int row = (y + dy) * w + x;
        for (int dx = -span; dx <= span; ++dx) {
            const float v = lvf(p, def, l, row + dx);
            mx += v * (float)dx;
            my += v * (float)dy;
        }
    }
    return atan2f(my, mx);
}

__device__ __forceinline__ float lv_bilinear_f(const PyrDev& p, bool def, int l, float x, float y) {
    const int w = p.w[l];
    const int xi = (int)floorf(x), yi = (int)floorf(y);
    const float fx = x - (float)xi, fy = y - (float)yi;
    const int i = yi * w + xi;
    const float a = lvf(p, def, l, i), b = lvf(p, def, l, i + 1);
    const float c = lvf(p, def, l, i + w), d = lvf(p, def, l, i + w + 1);
    return (a * (1.0f - fx) + b * fx) * (1.0f - fy) + (c * (1.0f - fx) + d * fx) * fy;
}

// polar sample k (ring-major) of the disc at (x, y) turned by theta
__device__ __forceinline__ float canon_sample(const PyrDev& p, bool def, int l, int x, int y, float ct, float st,
                                              int k) {
    const int ring = k / kCanonAngles, ang = k - ring * kCanonAngles;
    const float rho = (float)(ring + 2);
    float sa, ca;
    sincosf((float)ang * (6.283185307179586f / (float)kCanonAngles), &sa, &ca);
    // (ca, sa) turned by theta
    const float dx = rho * (ca * ct - sa * st), dy = rho * (sa * ct + ca * st);
    return lv_bilinear_f(p, def, l, (float)x + dx, (float)y + dy);
}

// One wave per reference patch: its orientation and normalised samples.
__global__ __launch_bounds__(64) void k_canon_ref(PyrDev p, int l, const int* __restrict__ centres, int npatch,
                                                  float* __restrict__ desc, float* __restrict__ theta) {
    const int q = blockIdx.x, lane = threadIdx.x;
    if (q >= npatch) return;
    const int x = centres[2 * q], y = centres[2 * q + 1];
    const int w = p.w[l], h = p.h[l];
    float* d = desc + (size_t)q * kCanonSamples;
    if (x < kCanonMargin || y < kCanonMargin || x >= w - kCanonMargin || y >= h - kCanonMargin) {
        for (int k = lane; k < kCanonSamples; k += 64) d[k] = 0.0f;
        if (lane == 0) theta[q] = 0.0f;
        return;
    }
    const float t = canon_theta(p, false, l, x, y);   // every lane the same value
    float ct, st;
    sincosf(t, &st, &ct);
    float v[(kCanonSamples + 63) / 64];
    float s = 0.0f, ss = 0.0f;
#pragma unroll
    for (int j = 0; j < (kCanonSamples + 63) / 64; ++j) {
        const int k = lane + 64 * j;
        v[j] = k < kCanonSamples ? canon_sample(p, false, l, x, y, ct, st, k) : 0.0f;
        s += v[j];
    }
    s = wave_sum(s);
    const float mean = s / (float)kCanonSamples;
#pragma unroll
    for (int j = 0; j < (kCanonSamples + 63) / 64; ++j) {
        const int k = lane + 64 * j;
        const float c = k < kCanonSamples ? v[j] - mean : 0.0f;
        ss += c * c;
    }
    ss = wave_sum(ss);
    const float inv = ss > 1e-12f ? 1.0f / sqrtf(ss) : 0.0f;
#pragma unroll
    for (int j = 0; j < (kCanonSamples + 63) / 64; ++j) {
        const int k = lane + 64 * j;
        if (k < kCanonSamples) d[k] = (v[j] - mean) * inv;
    }
    if (lane == 0) theta[q] = t;
}

// One thread per deformed position (x, y) of the level, the disc inside: the
// score against each patch's description, scores[pos * npatch + q]. The
// patch descriptions are zero-mean and unit-norm, so the correlation is the
// dot product over the position's own norm.
__global__ __launch_bounds__(256) void k_canon_scores(PyrDev p, int l, const float* __restrict__ desc, int npatch,
                                                      float* __restrict__ scores) {
    __shared__ float sd[kCanonMaxPatches * kCanonSamples];
    for (int i = threadIdx.x; i < npatch * kCanonSamples; i += blockDim.x) sd[i] = desc[i];
    __syncthreads();
    const int w = p.w[l], h = p.h[l];
    const int cw = w - 2 * kCanonMargin, chh = h - 2 * kCanonMargin;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (cw <= 0 || chh <= 0 || i >= cw * chh) return;
    const int x = kCanonMargin + i % cw, y = kCanonMargin + i / cw;
    const float t = canon_theta(p, true, l, x, y);
    float ct, st;
    sincosf(t, &st, &ct);
    float dot[kCanonMaxPatches];
#pragma unroll
    for (int q = 0; q < kCanonMaxPatches; ++q) dot[q] = 0.0f;
    float s = 0.0f, ss = 0.0f;
    for (int k = 0; k < kCanonSamples; ++k) {
        const float v = canon_sample(p, true, l, x, y, ct, st, k);
        s += v;
        ss += v * v;
#pragma unroll
        for (int q = 0; q < kCanonMaxPatches; ++q)
            if (q < npatch) dot[q] += v * sd[q * kCanonSamples + k];
    }
    const float var = ss - s * s / (float)kCanonSamples;
    const float inv = var > 1e-6f ? 1.0f / sqrtf(var) : 0.0f;
    float* o = scores + (size_t)i * npatch;
#pragma unroll
    for (int q = 0; q < kCanonMaxPatches; ++q)
        if (q < npatch) o[q] = inv > 0.0f ? dot[q] * inv : -2.0f;
}

// One block per patch: the best position (ties to the lower index), then the
// best one more than 2 px from it in either axis.
__global__ __launch_bounds__(256) void k_canon_best(PyrDev p, int l, const float* __restrict__ scores, int npatch,
                                                    const float* __restrict__ theta_ref, CanonOut* out) {
    __shared__ float bs[256];
    __shared__ int bi[256];
    const int q = blockIdx.x, tid = threadIdx.x;
    const int w = p.w[l], h = p.h[l];
    const int cw = w - 2 * kCanonMargin, chh = h - 2 * kCanonMargin;
    const int n = cw > 0 && chh > 0 ? cw * chh : 0;
    float best = -3.0f;
    int idx = -1;
    for (int i = tid; i < n; i += blockDim.x) {
        const float v = scores[(size_t)i * npatch + q];
        if (v > best) {
            best = v;
            idx = i;
        }
    }
    bs[tid] = best;
    bi[tid] = idx;
    __syncthreads();
    for (int m = 128; m > 0; m >>= 1) {
        if (tid < m) {
            const float a = bs[tid], b = bs[tid + m];
            const int ia = bi[tid], ib = bi[tid + m];
            if (b > a || (b == a && ib >= 0 && (ia < 0 || ib < ia))) {
                bs[tid] = b;
                bi[tid] = ib;
            }
        }
        __syncthreads();
    }
    const int ib = bi[0];
    const float b0 = bs[0];
    __syncthreads();
    const int bx = ib >= 0 ? ib % cw : 0, by = ib >= 0 ? ib / cw : 0;
    float sec = -3.0f;
    for (int i = tid; i < n; i += blockDim.x) {
        const int x = i % cw, y = i / cw;
        if (abs(x - bx) <= 2 && abs(y - by) <= 2) continue;
        sec = fmaxf(sec, scores[(size_t)i * npatch + q]);
    }
    bs[tid] = sec;
    __syncthreads();
    for (int m = 128; m > 0; m >>= 1) {
        if (tid < m) bs[tid] = fmaxf(bs[tid], bs[tid + m]);
        __syncthreads();
    }
    if (tid == 0) {
        CanonOut o = {};
        o.bx = ib >= 0 ? kCanonMargin + bx : -1;
        o.by = ib >= 0 ? kCanonMargin + by : -1;
        o.best = b0;
        o.second = bs[0];
        o.theta_ref = theta_ref[q];
        o.theta_def = ib >= 0 ? canon_theta(p, true, l, o.bx, o.by) : 0.0f;
        out[q] = o;
    }
}

int blocks_for(int64_t n, int bs) {
    const int64_t b = (n + bs - 1) / bs;
    return (int)(b < 65536 ? (b < 1 ? 1 : b) : 65536);
}

}  // namespace

hipError_t launch_prep_ref_u8(const uint8_t* img, const uint8_t* mask, int w, int h, float* I, float* gx,
                              float* gy, hipStream_t st) {
    hipLaunchKernelGGL(k_prep_ref<uint8_t>, dim3(blocks_for((int64_t)w * h, 256)), dim3(256), 0, st, img, mask,
                       w, h, I, gx, gy);
    return hipGetLastError();
}

hipError_t launch_prep_ref_f32(const float* img, int w, int h, float* I, float* gx, float* gy,
                               hipStream_t st) {
    hipLaunchKernelGGL(k_prep_ref<float>, dim3(blocks_for((int64_t)w * h, 256)), dim3(256), 0, st, img,
                       (const uint8_t*)nullptr, w, h, I, gx, gy);
    return hipGetLastError();
}

hipError_t launch_scan(const uint8_t* mask, int w, int h, int rect_x, int rect_y, int step, int gw, int gh,
                       int dim, uint8_t* skip, int* count, hipStream_t st, const int* dims) {
    const int n = gw * gh;
    if (n <= 0) return hipSuccess;
    if (mask)
        hipLaunchKernelGGL(k_scan_masked, dim3((n + kScanNodesPerGroup - 1) / kScanNodesPerGroup), dim3(256), 0, st,
                           mask, w, h, rect_x, rect_y, step, gw, gh, dim, dims, skip, count);
    else
        hipLaunchKernelGGL(k_scan, dim3((n + kScanBlock - 1) / kScanBlock), dim3(kScanBlock), 0, st, w, h, rect_x,
                           rect_y, step, gw, gh, dim, dims, skip, count);
    return hipGetLastError();
}

hipError_t launch_pool(const float* I, const float* gx, const float* gy, int w, int h, const uint8_t* skip,
                       int gw, int gh, int rect_x, int rect_y, int step, int dim, bool lm, PoolEntry* pool,
                       void* raw, hipStream_t st, const PoolOpts& o) {
    const int n = gw * gh;
    if (n <= 0) return hipSuccess;
    const GradRead G{gx, gy, o.gx16, o.gy16, o.bf16};
    PoolRaw* r = static_cast<PoolRaw*>(raw);
    if (o.mfma)
        hipLaunchKernelGGL(k_pool_sums<true>, dim3(n), dim3(64), 0, st, I, G, w, h, skip, gw, gh, rect_x, rect_y,
                           step, dim, o.dims, r);
    else
        hipLaunchKernelGGL(k_pool_sums<false>, dim3(n), dim3(64), 0, st, I, G, w, h, skip, gw, gh, rect_x, rect_y,
                           step, dim, o.dims, r);
    if (o.chol)
        hipLaunchKernelGGL(k_pool_tail<true>, dim3((n + 255) / 256), dim3(256), 0, st, r, n, dim, o.dims,
                           lm ? 1 : 0, pool, o.chol);
    else
        hipLaunchKernelGGL(k_pool_tail<false>, dim3((n + 255) / 256), dim3(256), 0, st, r, n, dim, o.dims,
                           lm ? 1 : 0, pool, nullptr);
    return hipGetLastError();
}

size_t pool_raw_bytes(int n) { return (size_t)n * sizeof(PoolRaw); }

__global__ void k_to_f16(const float* __restrict__ src, size_t n, uint16_t* __restrict__ dst) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = __builtin_bit_cast(uint16_t, (_Float16)src[i]);
}

// bf16 copy of a float plane (round to nearest even on the upper 16 bits)
__global__ void k_to_bf16(const float* __restrict__ src, size_t n, uint16_t* __restrict__ dst) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t u = __float_as_uint(src[i]);
    const uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
    dst[i] = (uint16_t)(r >> 16);
}

hipError_t launch_to_bf16(const float* src, size_t n, uint16_t* dst, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, n, dst);
    return hipGetLastError();
}

hipError_t launch_to_f16(const float* src, size_t n, uint16_t* dst, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_to_f16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, n, dst);
    return hipGetLastError();
}

// u8 -> FP32, four pixels per thread (exact: every u8 value is a float)
__global__ void k_u8_to_f32(const uint8_t* __restrict__ src, size_t n, float* __restrict__ dst) {
    const size_t i = 4 * ((size_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (i + 3 < n) {
        const uint32_t b = *reinterpret_cast<const uint32_t*>(src + i);
        const float4 v = make_float4((float)(b & 0xffu), (float)((b >> 8) & 0xffu), (float)((b >> 16) & 0xffu),
                                     (float)(b >> 24));
        *reinterpret_cast<float4*>(dst + i) = v;
    } else {
        for (size_t k = i; k < n; ++k) dst[k] = (float)src[k];
    }
}

hipError_t launch_u8_to_f32(const uint8_t* src, size_t n, float* dst, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const size_t threads = (n + 3) / 4;
    hipLaunchKernelGGL(k_u8_to_f32, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, src, n, dst);
    return hipGetLastError();
}

// Row-cubic coefficients of the Keys a = -0.5 kernel (image_processor.cpp:
// 188-195): a Keys row read at x = xi + s is sum_k w_k(s) p_k over the taps
// p_0..p_3 at xi - 1 .. xi + 2, the cubic c0 + c1 s + c2 s^2 + c3 s^3 with
//   c0 = p1, c1 = (p2 - p0) / 2, c2 = p0 - 5/2 p1 + 2 p2 - 1/2 p3,
//   c3 = -1/2 p0 + 3/2 p1 - 3/2 p2 + 1/2 p3.
// For u8 taps each c is a multiple of 1/2 with |c| <= 765, so both the FP32
// arithmetic here and the fp16 storage are exact. Entry (x, y) holds the
// cubics of the cell [x, x + 1) on rows y and y + 1, four fp16 each (16
// bytes, c0 lowest, row y first), so a 4 x 4 stencil is two 16-byte loads
// (entries (x, y - 1) and (x, y + 1)). Taps are clamped at the plane's edges
// (those cells lie inside the 4 px guard and are never read by the solver).
__device__ __forceinline__ uint2 row_cubic(const uint8_t* __restrict__ r, int x, int w) {
    const float p0 = r[x > 0 ? x - 1 : 0], p1 = r[x], p2 = r[x + 1 < w ? x + 1 : w - 1],
                p3 = r[x + 2 < w ? x + 2 : w - 1];
    const float c0 = p1, c1 = 0.5f * (p2 - p0), c2 = (p0 - 2.5f * p1) + (2.0f * p2 - 0.5f * p3),
                c3 = (1.5f * p1 - 0.5f * p0) + (0.5f * p3 - 1.5f * p2);
    const auto bits = [](float v) { return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)v); };
    return make_uint2(bits(c0) | bits(c1) << 16, bits(c2) | bits(c3) << 16);
}

__global__ void k_row_cubic_pairs(const uint8_t* __restrict__ img, int w, int h, uint4* __restrict__ out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= w) return;
    const int y = blockIdx.y;
    const uint2 a = row_cubic(img + (size_t)y * w, x, w);
    const uint2 b = row_cubic(img + (size_t)(y + 1 < h ? y + 1 : h - 1) * w, x, w);
    out[(size_t)y * w + x] = make_uint4(a.x, a.y, b.x, b.y);
}

hipError_t launch_row_cubic(const uint8_t* img, int w, int h, uint4* out, hipStream_t st) {
    if (w <= 0 || h <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_row_cubic_pairs, dim3((unsigned)((w + 255) / 256), (unsigned)h), dim3(256), 0, st, img, w,
                       h, out);
    return hipGetLastError();
}

hipError_t launch_pool_list(const float* I, const float* gx, const float* gy, int w, int h, const int* cxy,
                            int n, int dim, bool lm, PoolEntry* pool, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const GradRead G{gx, gy, nullptr, nullptr, false};
    hipLaunchKernelGGL(k_pool_list, dim3(n), dim3(64), 0, st, I, G, w, h, cxy, n, dim, lm ? 1 : 0, pool);
    return hipGetLastError();
}

hipError_t launch_downsample_u8(const uint8_t* src, int w, int h, float* dst, hipStream_t st) {
    hipLaunchKernelGGL(k_down<uint8_t>, dim3(blocks_for((int64_t)(w / 2) * (h / 2), 256)), dim3(256), 0, st,
                       src, w, h, dst);
    return hipGetLastError();
}

hipError_t launch_downsample_f32(const float* src, int w, int h, float* dst, hipStream_t st) {
    hipLaunchKernelGGL(k_down<float>, dim3(blocks_for((int64_t)(w / 2) * (h / 2), 256)), dim3(256), 0, st,
                       src, w, h, dst);
    return hipGetLastError();
}

size_t patch_scratch_bytes(int npatch) { return npatch > 0 ? (size_t)npatch * kPatchSplit * sizeof(Cand) : 0; }

hipError_t launch_patch_search(const PyrDev& p, const int* centres, int npatch, PatchOut* out, void* scratch,
                               hipStream_t st) {
    if (npatch <= 0) return hipSuccess;
    const int side = 2 * p.coarse_radius + 1;
    // split only a coarsest level of more candidates than the waves of one block take in a few rounds
    Cand* coarse = scratch && side * side > 4 * kPatchWaves ? (Cand*)scratch : nullptr;
    if (coarse)
        hipLaunchKernelGGL(k_patch_coarse, dim3(npatch, kPatchSplit), dim3(64 * kPatchWaves), 0, st, p, centres,
                           coarse);
    hipLaunchKernelGGL(k_patch, dim3(npatch), dim3(64 * kPatchWaves), 0, st, p, centres, (const Cand*)coarse, out);
    return hipGetLastError();
}

size_t canon_scores_floats(int wl, int hl, int npatch) {
    const int cw = wl - 2 * kCanonMargin, ch = hl - 2 * kCanonMargin;
    return cw > 0 && ch > 0 ? (size_t)cw * ch * npatch : 0;
}

hipError_t launch_canon_match(const PyrDev& p, int l, const int* centres_l, int npatch, float* desc,
                              float* theta_ref, float* scores, CanonOut* out, hipStream_t st) {
    if (npatch <= 0 || npatch > kCanonMaxPatches || l < 0 || l > p.levels) return hipErrorInvalidValue;
    const size_t n = canon_scores_floats(p.w[l], p.h[l], npatch) / npatch;
    if (n == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_canon_ref, dim3(npatch), dim3(64), 0, st, p, l, centres_l, npatch, desc, theta_ref);
    hipLaunchKernelGGL(k_canon_scores, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, l, desc, npatch,
                       scores);
    hipLaunchKernelGGL(k_canon_best, dim3(npatch), dim3(256), 0, st, p, l, scores, npatch, theta_ref, out);
    return hipGetLastError();
}

}  // namespace sdic
