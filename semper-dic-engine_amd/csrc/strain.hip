// Post-solve kernels, one thread per grid node:
//   k_outlier  one pass of the normalized median test (Westerweel & Scarano)
//              on u and v over the 5x5 neighbourhood, reading only the
//              previous pass's flags (strain_calculator.cpp:164-224)
//   k_vsg      VSG plane fit in FP64 over the circular window, 90% fill,
//              1-norm rcond, Green-Lagrange strain (strain_calculator.cpp:27-158)
#include <hip/hip_runtime.h>

#include <utility>

#include "common.hpp"
#include "solver.hpp"

#pragma clang fp contract(off)

namespace sdic {

namespace {

constexpr int kNb = 24;   // 5x5 minus the centre

// Batcher's merge-exchange network (Knuth, TAOCP vol. 3, 5.2.2, Algorithm M)
// for kNb keys, as a list of compare-exchange pairs fixed at compile time.
struct SortNet {
    int n = 0;
    uint8_t a[160] = {}, b[160] = {};
};
constexpr SortNet merge_exchange(int keys) {
    SortNet s;
    int t = 0;
    while ((1 << t) < keys) ++t;
    for (int p = 1 << (t - 1); p > 0; p >>= 1) {
        int q = 1 << (t - 1), r = 0, d = p;
        while (d > 0) {
            for (int i = 0; i + d < keys; ++i)
                if ((i & p) == r) {
                    s.a[s.n] = (uint8_t)i;
                    s.b[s.n] = (uint8_t)(i + d);
                    ++s.n;
                }
            d = q - p;
            q >>= 1;
            r = p;
        }
    }
    return s;
}
constexpr SortNet kNet = merge_exchange(kNb);
static_assert(kNet.n <= 160, "network larger than its table");

// One compare-exchange of the network: its pair is a template argument, so the
// keys stay in registers.
template <int C>
__device__ __forceinline__ void cmp_exchange(float (&v)[kNb]) {
    constexpr int i = kNet.a[C], j = kNet.b[C];
    const float lo = fminf(v[i], v[j]), hi = fmaxf(v[i], v[j]);
    v[i] = lo;
    v[j] = hi;
}
template <int... C>
__device__ __forceinline__ void sort_net(float (&v)[kNb], std::integer_sequence<int, C...>) {
    (cmp_exchange<C>(v), ...);
}

// The median of the `valid` entries of v (median_inplace: the middle value,
// or the mean of the middle pair). The invalid entries become +inf and the
// network sorts ascending, so the k valid values lead in order; an order
// statistic is the same value whichever selection finds it.
__device__ __forceinline__ float median_masked(float (&v)[kNb], uint32_t valid, int k) {
#pragma unroll
    for (int i = 0; i < kNb; ++i)
        if (!((valid >> i) & 1)) v[i] = __builtin_inff();
    sort_net(v, std::make_integer_sequence<int, kNet.n>{});
    const int r_hi = k / 2, r_lo = (k - 1) / 2;
    float hi = 0.0f, lo = 0.0f;
#pragma unroll
    for (int i = 0; i < kNb; ++i) {
        hi = (i == r_hi) ? v[i] : hi;
        lo = (i == r_lo) ? v[i] : lo;
    }
    return (k & 1) ? hi : 0.5f * (lo + hi);
}

__device__ __forceinline__ float norm_residual(float c, const float (&nb)[kNb], uint32_t valid, int k) {
    float w[kNb];
#pragma unroll
    for (int i = 0; i < kNb; ++i) w[i] = nb[i];
    const float med = median_masked(w, valid, k);
#pragma unroll
    for (int i = 0; i < kNb; ++i) w[i] = fabsf(nb[i] - med);
    const float spread = median_masked(w, valid, k);
    return fabsf(c - med) / (spread + kOutlierEpsPx);
}

__global__ void k_outlier(int gw, int gh, const float* __restrict__ u, const float* __restrict__ v,
                          const uint8_t* __restrict__ ok_in, uint8_t* __restrict__ ok_out,
                          const int* __restrict__ prev_fail, int* nfail) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= gw * gh) return;
    // a pass after one that rejected nothing sees the same flags: it is the identity
    if (prev_fail && *prev_fail == 0) {
        ok_out[idx] = ok_in[idx];
        return;
    }
    if (!ok_in[idx]) {
        ok_out[idx] = 0;
        return;
    }
    const int x = idx % gw, y = idx / gw;
    // the 24 neighbours, loaded unconditionally (a slot off the grid reads the
    // node itself) so that the keys stay in registers; `valid` marks the ones
    // that count
    float nu[kNb], nv[kNb];
    uint32_t valid = 0;
#pragma unroll
    for (int s = 0; s < kNb; ++s) {
        const int t = s < kNb / 2 ? s : s + 1;   // the 5x5 window, row-major, without its centre
        const int nx = x + t % 5 - 2, ny = y + t / 5 - 2;
        const bool in = nx >= 0 && nx < gw && ny >= 0 && ny < gh;
        const int j = in ? ny * gw + nx : idx;
        nu[s] = u[j];
        nv[s] = v[j];
        if (in && ok_in[j]) valid |= (1u << s);
    }
    const int k = __builtin_popcount(valid);
    bool fail = false;
    if (k >= kOutlierMinNeighbours)
        fail = norm_residual(u[idx], nu, valid, k) > kOutlierThreshold ||
               norm_residual(v[idx], nv, valid, k) > kOutlierThreshold;
    ok_out[idx] = fail ? 0 : 1;
    if (fail) atomicAdd(nfail, 1);
}

__global__ void k_vsg(int gw, int gh, int step, int window, int expected, const float* __restrict__ u,
                      const float* __restrict__ v, const uint8_t* __restrict__ valid, float* __restrict__ exx,
                      float* __restrict__ eyy, float* __restrict__ exy) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= gw * gh) return;
    exx[idx] = kStrainUninit;
    eyy[idx] = kStrainUninit;
    exy[idx] = kStrainUninit;
    if (!valid[idx]) return;
    const int x = idx % gw, y = idx / gw;
    const float radius = window / 2.0f;
    const float radius_sq = radius * radius;
    const int grid_rad = (int)ceilf(radius / step);
    const double r2 = (double)radius_sq + kVsgRadiusTiny;
    double A[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, bu[3] = {0, 0, 0}, bv[3] = {0, 0, 0};
    int np = 0;
    for (int dy = -grid_rad; dy <= grid_rad; ++dy)
        for (int dx = -grid_rad; dx <= grid_rad; ++dx) {
            const int nx = x + dx, ny = y + dy;
            if (nx < 0 || nx >= gw || ny < 0 || ny >= gh) continue;
            const int nidx = ny * gw + nx;
            if (!valid[nidx]) continue;
            const double px = (double)(dx * step), py = (double)(dy * step);
            if (px * px + py * py > r2) continue;
            const double a[3] = {1.0, px, py};
            for (int r = 0; r < 3; ++r) {
                for (int c = 0; c < 3; ++c) A[r * 3 + c] += a[r] * a[c];
                bu[r] += a[r] * (double)u[nidx];
                bv[r] += a[r] * (double)v[nidx];
            }
            np++;
        }
    const double fill = (double)np / (double)expected;
    if (!(fill >= 0.90 && np >= 3)) return;
    double anorm = 0.0;
    for (int c = 0; c < 3; ++c) anorm = fmax(anorm, fabs(A[c]) + fabs(A[3 + c]) + fabs(A[6 + c]));
    auto cof = [&](int i, int j) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        return A[i1 * 3 + j1] * A[i2 * 3 + j2] - A[i1 * 3 + j2] * A[i2 * 3 + j1];
    };
    const double det = cof(0, 0) * A[0] + cof(1, 0) * A[3] + cof(2, 0) * A[6];
    if (!(fabs(det) > 1e-12)) return;
    double Ai[9];
    const double inv = 1.0 / det;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Ai[i * 3 + j] = cof(j, i) * inv;
    double inorm = 0.0;
    for (int c = 0; c < 3; ++c) inorm = fmax(inorm, fabs(Ai[c]) + fabs(Ai[3 + c]) + fabs(Ai[6 + c]));
    const double rcond = (anorm * inorm > 0.0) ? 1.0 / (anorm * inorm) : 0.0;
    if (rcond < 1e-12) return;
    double cu[3], cv[3];
    for (int r = 0; r < 3; ++r) {
        cu[r] = Ai[r * 3] * bu[0] + Ai[r * 3 + 1] * bu[1] + Ai[r * 3 + 2] * bu[2];
        cv[r] = Ai[r * 3] * bv[0] + Ai[r * 3 + 1] * bv[1] + Ai[r * 3 + 2] * bv[2];
    }
    const double ux = cu[1], uy = cu[2], vx = cv[1], vy = cv[2];
    exx[idx] = (float)(0.5 * (2.0 * ux + ux * ux + vx * vx));
    eyy[idx] = (float)(0.5 * (2.0 * vy + uy * uy + vy * vy));
    exy[idx] = (float)(0.5 * (uy + vx + ux * uy + vx * vy));
}

// ---- row-major pack (full_field_solver_stats.cpp:21-91) ---------------------
// A node that passed the median test is written when its strain succeeded
// (exx > kStrainFail) and dropped (slot 20) otherwise. Packing stops at the
// first such node that does not fit: with pos = the number of written nodes
// before a node and cap = out_capacity / 8, a node is written iff pos < cap,
// and a dropped node is counted iff pos < cap.
constexpr int kPackChunk = 1024;   // nodes per block (256 threads x 4)

__device__ __forceinline__ void pack_flags(int i, int n, const uint8_t* ok, const float* exx, int& wr, int& dr) {
    wr = 0;
    dr = 0;
    if (i < n && ok[i]) {
        if (exx[i] > kStrainFail) wr = 1;
        else dr = 1;
    }
}

__global__ __launch_bounds__(256) void k_pack_count(int n, const uint8_t* __restrict__ ok,
                                                    const float* __restrict__ exx, int* __restrict__ block_w) {
    __shared__ int part[4];
    const int base = blockIdx.x * kPackChunk;
    int c = 0;
    for (int k = threadIdx.x; k < kPackChunk; k += 256) {
        int wr, dr;
        pack_flags(base + k, n, ok, exx, wr, dr);
        c += wr;
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) c += __shfl_xor(c, m, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) block_w[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// exclusive scan of the block counts in place, total into *total (one block)
__global__ __launch_bounds__(1024) void k_pack_scan(int nb, int* __restrict__ block_w, int* __restrict__ total) {
    __shared__ int carry;
    __shared__ int wsum[16];
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int base = 0; base < nb; base += 1024) {
        const int i = base + threadIdx.x;
        const int v = i < nb ? block_w[i] : 0;
        int incl = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int t = __shfl_up(incl, d, 64);
            if (lane >= d) incl += t;
        }
        if (lane == 63) wsum[wid] = incl;
        __syncthreads();
        int before = carry;
        for (int k = 0; k < wid; ++k) before += wsum[k];
        if (i < nb) block_w[i] = before + incl - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry = before + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(256) void k_pack_write(int n, int gw, int rect_x, int rect_y, int step, int cap,
                                                    const uint8_t* __restrict__ ok, const float* __restrict__ u,
                                                    const float* __restrict__ v, const float* __restrict__ exx,
                                                    const float* __restrict__ eyy, const float* __restrict__ exy,
                                                    const float* __restrict__ corr,
                                                    const int* __restrict__ block_w, float* __restrict__ out,
                                                    int* __restrict__ dropped) {
    __shared__ int wsum[4];
    __shared__ int running;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int base = blockIdx.x * kPackChunk;
    if (threadIdx.x == 0) running = block_w[blockIdx.x];
    __syncthreads();
    int drops = 0;
    // four rounds of 256 consecutive nodes, in row-major order
    for (int r = 0; r < kPackChunk / 256; ++r) {
        const int i = base + r * 256 + threadIdx.x;
        int wr, dr;
        pack_flags(i, n, ok, exx, wr, dr);
        const unsigned long long m = __ballot(wr);
        const int in_wave = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wid] = __popcll(m);
        __syncthreads();
        int pos = running + in_wave;
        for (int k = 0; k < wid; ++k) pos += wsum[k];
        if (wr && pos < cap) {
            float* o = out + (size_t)pos * 8;
            o[0] = (float)(rect_x + (i % gw) * step);
            o[1] = (float)(rect_y + (i / gw) * step);
            o[2] = u[i];
            o[3] = v[i];
            o[4] = exx[i];
            o[5] = eyy[i];
            o[6] = exy[i];
            o[7] = corr[i];
        }
        if (dr && pos < cap) drops++;
        __syncthreads();
        if (threadIdx.x == 0) running += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) drops += __shfl_xor(drops, m, 64);
    if (lane == 0 && drops) atomicAdd(dropped, drops);
}

}  // namespace

int pack_blocks(int n) { return (n + kPackChunk - 1) / kPackChunk; }

hipError_t launch_pack(int gw, int gh, int rect_x, int rect_y, int step, int cap_points, const uint8_t* ok,
                       const float* u, const float* v, const float* exx, const float* eyy, const float* exy,
                       const float* corr, int* block_w, int* counts, float* out, hipStream_t st) {
    const int n = gw * gh;
    const int nb = pack_blocks(n);
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_count, dim3(nb), dim3(256), 0, st, n, ok, exx, block_w);
    hipLaunchKernelGGL(k_pack_scan, dim3(1), dim3(1024), 0, st, nb, block_w, counts);
    hipLaunchKernelGGL(k_pack_write, dim3(nb), dim3(256), 0, st, n, gw, rect_x, rect_y, step, cap_points, ok, u, v,
                       exx, eyy, exy, corr, block_w, out, counts + 1);
    return hipGetLastError();
}

hipError_t launch_outlier_pass(int gw, int gh, const float* u, const float* v, const uint8_t* ok_in,
                               uint8_t* ok_out, const int* prev_fail, int* nfail, hipStream_t st) {
    const int n = gw * gh;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_outlier, dim3((n + 255) / 256), dim3(256), 0, st, gw, gh, u, v, ok_in, ok_out, prev_fail,
                       nfail);
    return hipGetLastError();
}

hipError_t launch_vsg(int gw, int gh, int step, int window, int expected, const float* u, const float* v,
                      const uint8_t* valid, float* exx, float* eyy, float* exy, hipStream_t st) {
    const int n = gw * gh;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_vsg, dim3((n + 255) / 256), dim3(256), 0, st, gw, gh, step, window, expected, u, v,
                       valid, exx, eyy, exy);
    return hipGetLastError();
}

}  // namespace sdic
