// Per-node bookkeeping kernels of the device grid (the ResultGrid as planes).
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "solver.hpp"
#include "wave.hpp"

namespace sdic {

namespace {

__global__ void k_grid_init(GridDev g, const uint8_t* __restrict__ skip, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    g.u[i] = g.v[i] = g.ux[i] = g.uy[i] = g.vx[i] = g.vy[i] = 0.0f;
    g.corr[i] = kCorrInvalid;
    g.solved[i] = skip[i];
    g.tag[i] = kMeshNone;
    g.iters[i] = 0;
    g.simplex[i] = 0;
}

__global__ void k_apply(GridDev g, const int* __restrict__ nodes, const SolveOut* __restrict__ res, int n, int tag,
                        int accepted_only) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = nodes[k];
    if (i < 0) return;
    const SolveOut o = res[k];
    if (!o.init) return;
    const bool acc = o.status == 0 && o.corr <= kCorrAccept;
    g.iters[i] = o.iters;
    g.simplex[i] = o.simplex ? 1 : 0;
    if (!accepted_only || acc) {
        g.u[i] = o.p[0];
        g.v[i] = o.p[1];
        g.ux[i] = o.p[2];
        g.uy[i] = o.p[3];
        g.vx[i] = o.p[4];
        g.vy[i] = o.p[5];
        g.corr[i] = o.corr;
        g.solved[i] = 1;
        if (tag >= 0) g.tag[i] = (uint8_t)tag;
    } else {
        g.corr[i] = kCorrInvalid;
    }
}

__global__ void k_valid(GridDev g, uint8_t* __restrict__ valid, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    valid[i] = (g.solved[i] && g.corr[i] >= 0.0f) ? 1 : 0;
}

// Nodes Path B can still claim (solved == 0), one atomic per wave.
__global__ void k_count_open(GridDev g, int n, int* __restrict__ count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool open = i < n && !g.solved[i];
    const unsigned long long m = __ballot(open);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(count, __popcll(m));
}

__global__ void k_zero1(int* p) { *p = 0; }

// Anchor patches of the incremental session: entry e = (slot, x0, y0) copies
// the side x side block of the frame whose top-left pixel is (x0, y0) into
// store slot `slot` (row-major), one workgroup per entry.
__global__ void k_cut_patches(const uint8_t* __restrict__ frame, int w, const int32_t* __restrict__ entries,
                              int side, uint8_t* __restrict__ store) {
    const int e = blockIdx.x;
    const int slot = entries[3 * e], x0 = entries[3 * e + 1], y0 = entries[3 * e + 2];
    const int n = side * side;
    uint8_t* dst = store + (size_t)slot * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int y = i / side, x = i - y * side;
        dst[i] = frame[(size_t)(y0 + y) * w + x0 + x];
    }
}

// The matching canvas: patch k of the list (store slot slots[k]) at column
// k % cols, row k / cols of side x side cells on a w-wide canvas whose other
// bytes are zero (cleared by k_fill_u8 first).
__global__ void k_place_patches(const uint8_t* __restrict__ store, int side, const int32_t* __restrict__ slots,
                                int cols, int w, uint8_t* __restrict__ canvas) {
    const int k = blockIdx.x;
    const int n = side * side;
    const uint8_t* src = store + (size_t)slots[k] * n;
    const int ox = (k % cols) * side, oy = (k / cols) * side;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int y = i / side, x = i - y * side;
        canvas[(size_t)(oy + y) * w + ox + x] = src[i];
    }
}

__global__ void k_fill_u8(uint8_t* p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_zero_n(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// wave_sum / wave_sum_i over one wave per 64 inputs; every lane writes its result
__global__ void k_wave_sum_probe(const float* __restrict__ in, int nw, float* __restrict__ out,
                                 int* __restrict__ iout) {
    const int wv = blockIdx.x;
    if (wv >= nw) return;
    const int i = wv * 64 + threadIdx.x;
    out[i] = wave_sum(in[i]);
    iout[i] = wave_sum_i((int)in[i]);
}

}  // namespace

hipError_t launch_wave_sum_probe(const float* in, int nw, float* out, int* iout, hipStream_t st) {
    if (nw <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_wave_sum_probe, dim3(nw), dim3(64), 0, st, in, nw, out, iout);
    return hipGetLastError();
}

hipError_t launch_zero(int* p, int n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_n, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_count_open(GridDev g, int n, int* count, hipStream_t st) {
    hipLaunchKernelGGL(k_zero1, dim3(1), dim3(1), 0, st, count);
    if (n > 0) hipLaunchKernelGGL(k_count_open, dim3((n + 255) / 256), dim3(256), 0, st, g, n, count);
    return hipGetLastError();
}

hipError_t launch_grid_init(GridDev g, const uint8_t* skip, int n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_grid_init, dim3((n + 255) / 256), dim3(256), 0, st, g, skip, n);
    return hipGetLastError();
}

hipError_t launch_apply(GridDev g, const int* nodes, const SolveOut* res, int n, int tag, bool accepted_only,
                        hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_apply, dim3((n + 255) / 256), dim3(256), 0, st, g, nodes, res, n, tag,
                       accepted_only ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_cut_patches(const uint8_t* frame, int w, const int32_t* entries, int n, int side, uint8_t* store,
                              hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cut_patches, dim3(n), dim3(256), 0, st, frame, w, entries, side, store);
    return hipGetLastError();
}

hipError_t launch_place_patches(const uint8_t* store, int side, const int32_t* slots, int n, int cols, int w,
                                int rows, uint8_t* canvas, hipStream_t st) {
    const size_t npx = (size_t)w * rows * side;
    if (npx > 0) hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0, st, canvas, npx);
    if (n > 0) hipLaunchKernelGGL(k_place_patches, dim3(n), dim3(256), 0, st, store, side, slots, cols, w, canvas);
    return hipGetLastError();
}

hipError_t launch_valid(GridDev g, uint8_t* valid, int n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_valid, dim3((n + 255) / 256), dim3(256), 0, st, g, valid, n);
    return hipGetLastError();
}

}  // namespace sdic
