// Post-solve kernels, one thread per grid node:
//   k_outlier  one pass of the normalized median test (Westerweel & Scarano)
//              on u and v over the 5x5 neighbourhood, reading only the
//              previous pass's flags (strain_calculator.cpp:164-224)
//   k_vsg      VSG plane fit in FP64 over the circular window, 90% fill,
//              1-norm rcond, Green-Lagrange strain (strain_calculator.cpp:27-158)
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "solver.hpp"

#pragma clang fp contract(off)

namespace sdic {

namespace {

constexpr int kNb = 24;   // 5x5 minus the centre

// Median of the `valid` entries of v[0..kNb), by rank (ties by index), as
// median_inplace: the middle value, or the mean of the middle pair.
__device__ __forceinline__ float median_masked(const float (&v)[kNb], uint32_t valid, int k) {
    const int r_hi = k / 2, r_lo = (k - 1) / 2;
    float hi = 0.0f, lo = 0.0f;
#pragma unroll
    for (int i = 0; i < kNb; ++i) {
        if (!((valid >> i) & 1)) continue;
        int r = 0;
#pragma unroll
        for (int j = 0; j < kNb; ++j)
            r += ((valid >> j) & 1) && (v[j] < v[i] || (v[j] == v[i] && j < i));
        if (r == r_hi) hi = v[i];
        if (r == r_lo) lo = v[i];
    }
    return (k & 1) ? hi : 0.5f * (lo + hi);
}

__device__ __forceinline__ float norm_residual(float c, const float (&nb)[kNb], uint32_t valid, int k) {
    const float med = median_masked(nb, valid, k);
    float sc[kNb];
#pragma unroll
    for (int i = 0; i < kNb; ++i) sc[i] = fabsf(nb[i] - med);
    const float spread = median_masked(sc, valid, k);
    return fabsf(c - med) / (spread + kOutlierEpsPx);
}

__global__ void k_outlier(int gw, int gh, const float* __restrict__ u, const float* __restrict__ v,
                          const uint8_t* __restrict__ ok_in, uint8_t* __restrict__ ok_out, int* nfail) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= gw * gh) return;
    if (!ok_in[idx]) {
        ok_out[idx] = 0;
        return;
    }
    const int x = idx % gw, y = idx / gw;
    float nu[kNb], nv[kNb];
    uint32_t valid = 0;
    int k = 0, slot = 0;
#pragma unroll
    for (int dy = -2; dy <= 2; ++dy)
#pragma unroll
        for (int dx = -2; dx <= 2; ++dx) {
            if (dx == 0 && dy == 0) continue;
            const int nx = x + dx, ny = y + dy;
            nu[slot] = 0.0f;
            nv[slot] = 0.0f;
            if (nx >= 0 && nx < gw && ny >= 0 && ny < gh && ok_in[ny * gw + nx]) {
                nu[slot] = u[ny * gw + nx];
                nv[slot] = v[ny * gw + nx];
                valid |= (1u << slot);
                k++;
            }
            slot++;
        }
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

}  // namespace

hipError_t launch_outlier_pass(int gw, int gh, const float* u, const float* v, const uint8_t* ok_in,
                               uint8_t* ok_out, int* nfail, hipStream_t st) {
    const int n = gw * gh;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_outlier, dim3((n + 255) / 256), dim3(256), 0, st, gw, gh, u, v, ok_in, ok_out, nfail);
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
