// Subset-solver instantiations: bicubic 4x4 interpolation with the fast path's
// interior samples read from the row-cubic coefficient plane (the full solver
// and the guard-band samples read the FP32 deformed plane).
#include "solver_impl.hpp"

namespace sdic {

hipError_t launch_solve_cf32(const SolveArgs& a, int ppl, hipStream_t st, int phases) {
    if (a.nitems <= 0) return hipSuccess;
    const int wpb = a.wpb == 2 || a.wpb == 4 ? a.wpb : 1;
    const dim3 block(64 * kSlowWaves), fast_block(64 * wpb), fast_grid((a.nitems + wpb - 1) / wpb),
        slow_grid(a.nitems < 2048 ? a.nitems : 2048);
    switch (ppl) {
#define SDIC_CASE(P)                                                                    \
    case P:                                                                             \
        if (phases & 1) hipLaunchKernelGGL((k_fast<P, false, float, true>), fast_grid, fast_block, 0, st, a); \
        if (phases & 2) hipLaunchKernelGGL((k_slow<P, false, float, kSlowWaves>), slow_grid, block, 0, st, a); \
        break;
        SDIC_CASE(4)
        SDIC_CASE(8)
        SDIC_CASE(12)
        SDIC_CASE(16)
        SDIC_CASE(24)
        SDIC_CASE(32)
        SDIC_CASE(48)
        SDIC_CASE(64)
#undef SDIC_CASE
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// The fast path's Keys 4x4 sampling forms at given points, for tests: form 0
// the tap weights over the FP32 plane, 1 the row-cubic coefficient plane,
// 2 the tap weights over the u8 plane (the caller keeps every point inside
// the solver's 4 px guard).
__global__ void k_interp_probe(const float* __restrict__ f32, const uint4* __restrict__ coef,
                               const uint8_t* __restrict__ u8, int w, int h, int form, int n,
                               const float* __restrict__ xs, const float* __restrict__ ys, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = xs[i], y = ys[i];
    const int npx = w * h;
    if (form == 1) {
        const auto rs = __builtin_amdgcn_make_buffer_rsrc((void*)coef, (short)0, npx * 16, 0x00020000);
        out[i] = interp_coef(rs, w, x, y);
    } else if (form == 0) {
        const auto rs = __builtin_amdgcn_make_buffer_rsrc((void*)f32, (short)0, npx * 4, 0x00020000);
        out[i] = interp_buf<false, float>(rs, w, x, y);
    } else {
        const auto rs = __builtin_amdgcn_make_buffer_rsrc((void*)u8, (short)0, npx, 0x00020000);
        out[i] = interp_buf<false, uint8_t>(rs, w, x, y);
    }
}

hipError_t launch_interp_probe(const float* f32, const uint4* coef, const uint8_t* u8, int w, int h, int form,
                               int n, const float* xs, const float* ys, float* out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_interp_probe, dim3((n + 255) / 256), dim3(256), 0, st, f32, coef, u8, w, h, form, n, xs, ys,
                       out);
    return hipGetLastError();
}

}  // namespace sdic
