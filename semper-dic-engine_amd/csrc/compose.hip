// FP64 composition of an incremental session's totals on the device (no
// reference counterpart: the reference adds each increment to a float total,
// incremental_session.cpp:411-412).
//
// One thread per node of the session's frame-0 grid. A node still followed
// samples the increment field (the last run's node planes and median-test
// flags, on the device) at its current position x0 + U, exactly as the host's
// sample_increment does (incremental_math.cpp:61-123: the node itself on an
// exact hit, bilinear over four valid corners, else the first-order
// expansion from the nearest valid node of the 4 x 4 block within 1.5 cells),
// in the same float operations; the increment is then added to the total in
// double and the reported total is that sum rounded once -- or, for a session
// with float totals (no double totals given), added to the float total as the
// reference does.
#include <hip/hip_runtime.h>

#include "compose.hpp"

#pragma clang fp contract(off)

namespace sdic {

namespace {

constexpr float kTaylorCells = 1.5f;

// std::max's order: the first argument unless it is less than the second
__device__ __forceinline__ float max_of(float a, float b) { return a < b ? b : a; }

__device__ __forceinline__ bool valid_at(const ComposeField& f, int i, int j) {
    return i >= 0 && i < f.gw && j >= 0 && j < f.gh && f.med[(size_t)j * f.gw + i] != 0;
}

__global__ void k_sample_compose64(ComposeField f, ComposeTotals t, ComposeOut o) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= t.gw * t.gh) return;
    o.flags[k] = 0;
    if (!t.alive[k]) return;
    const float x = (float)(t.rect_x + (k % t.gw) * t.step) + t.uf[k];
    const float y = (float)(t.rect_y + (k / t.gw) * t.step) + t.vf[k];
    const size_t NN = (size_t)f.gw * f.gh;
    const float* U = f.planes;
    const float* V = U + NN;
    const float* UX = U + 2 * NN;
    const float* UY = U + 3 * NN;
    const float* VX = U + 4 * NN;
    const float* VY = U + 5 * NN;
    const float* C = U + 6 * NN;
    const float gx = (x - (float)f.rect_x) / (float)f.step, gy = (y - (float)f.rect_y) / (float)f.step;
    const int i0 = (int)floorf(gx), j0 = (int)floorf(gy);
    const float fx = gx - (float)i0, fy = gy - (float)j0;
    float su = 0.0f, sv = 0.0f, sc = 0.0f;
    bool ok = false, taylor = false;
    if (fx == 0.0f && fy == 0.0f && valid_at(f, i0, j0)) {
        const size_t a = (size_t)j0 * f.gw + i0;
        su = U[a];
        sv = V[a];
        sc = C[a];
        ok = true;
    } else if (valid_at(f, i0, j0) && valid_at(f, i0 + 1, j0) && valid_at(f, i0, j0 + 1) &&
               valid_at(f, i0 + 1, j0 + 1)) {
        const size_t a = (size_t)j0 * f.gw + i0, b = a + 1, c = a + f.gw, d = c + 1;
        const float wa = (1 - fx) * (1 - fy), wb = fx * (1 - fy), wc = (1 - fx) * fy, wd = fx * fy;
        su = wa * U[a] + wb * U[b] + wc * U[c] + wd * U[d];
        sv = wa * V[a] + wb * V[b] + wc * V[c] + wd * V[d];
        sc = max_of(max_of(C[a], C[b]), max_of(C[c], C[d]));
        ok = true;
    } else {
        const float max_d = kTaylorCells * (float)f.step;
        float best = __builtin_inff();
        int bi = -1, bj = -1;
        for (int j = j0 - 1; j <= j0 + 2; ++j)
            for (int i = i0 - 1; i <= i0 + 2; ++i) {
                if (!valid_at(f, i, j)) continue;
                const float dx = x - (float)(f.rect_x + i * f.step), dy = y - (float)(f.rect_y + j * f.step);
                const float d2 = dx * dx + dy * dy;
                if (d2 < best) {
                    best = d2;
                    bi = i;
                    bj = j;
                }
            }
        if (bi >= 0 && !(best > max_d * max_d)) {
            const size_t a = (size_t)bj * f.gw + bi;
            const float dx = x - (float)(f.rect_x + bi * f.step), dy = y - (float)(f.rect_y + bj * f.step);
            su = U[a] + UX[a] * dx + UY[a] * dy;
            sv = V[a] + VX[a] * dx + VY[a] * dy;
            sc = C[a];
            ok = true;
            taylor = true;
        }
    }
    if (!ok) return;
    if (t.u64) {
        const double U64 = t.u64[k] + (double)su, V64 = t.v64[k] + (double)sv;
        o.u64[k] = U64;
        o.v64[k] = V64;
        o.uf[k] = (float)U64;
        o.vf[k] = (float)V64;
    } else {
        // the reference's float totals: the increment added in float
        o.uf[k] = t.uf[k] + su;
        o.vf[k] = t.vf[k] + sv;
    }
    o.corr[k] = sc;
    o.flags[k] = (uint8_t)(1 | (taylor ? 2 : 0));
}

}  // namespace

hipError_t launch_sample_compose64(const ComposeField& f, const ComposeTotals& t, const ComposeOut& o,
                                   hipStream_t st) {
    const int n = t.gw * t.gh;
    if (n <= 0) return hipSuccess;
    if (f.gw <= 0 || f.gh <= 0 || f.step <= 0 || !f.planes || !f.med) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_sample_compose64, dim3((n + 255) / 256), dim3(256), 0, st, f, t, o);
    return hipGetLastError();
}

}  // namespace sdic
