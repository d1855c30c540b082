// Dispatch of the subset solver over its instantiations, and the pixels-per-lane
// class of a subset size.
#include "solver.hpp"

#include <cmath>

#include "wave.hpp"

namespace sdic {

hipError_t launch_solve_bu8(const SolveArgs& a, int ppl, hipStream_t st, int phases);
hipError_t launch_solve_ku8(const SolveArgs& a, int ppl, hipStream_t st, int phases);
hipError_t launch_solve_bf32(const SolveArgs& a, int ppl, hipStream_t st, int phases);
hipError_t launch_solve_kf32(const SolveArgs& a, int ppl, hipStream_t st, int phases);
hipError_t launch_solve_cf32(const SolveArgs& a, int ppl, hipStream_t st, int phases);

int ppl_for(int dim) {
    if (dim < 3 || dim > kMaxSubset) return -1;
    const int p = (dim * dim + 63) / 64;
    static const int cls[] = {4, 8, 12, 16, 24, 32, 48, 64};
    for (int c : cls)
        if (p <= c) return c;
    return -1;
}

__global__ void k_zero_int(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// one wave per counter: each lane adds kStatSlots / 64 slots, then the wave
// sum (integer sums, so the order does not matter)
__global__ void k_sum_stats(const int* __restrict__ slots, int* __restrict__ out) {
    const int c = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int s = 0;
    for (int k = lane; k < kStatSlots; k += 64) s += slots[k * kStatStride + c];
    s = wave_sum_i(s);
    if (lane == 0) out[c] = s;
}

hipError_t launch_sum_stats(const int* slots, int* out, hipStream_t st) {
    hipLaunchKernelGGL(k_sum_stats, dim3(1), dim3(5 * 64), 0, st, slots, out);
    return hipGetLastError();
}

// The least float x with sqrtf(x) >= tol (correctly rounded, as on the
// device): sqrtf is monotone, so sqrtf(x) < tol exactly when x < the result.
static float stop_threshold(float tol) {
    if (!(tol > 0.0f)) return tol == 0.0f ? 0.0f : tol;   // 0: never below; NaN: never
    if (std::isinf(tol)) return tol;
    float x = (float)((double)tol * (double)tol);
    while (x > 0.0f && std::sqrt(std::nextafter(x, 0.0f)) >= tol) x = std::nextafter(x, 0.0f);
    while (std::sqrt(x) < tol) x = std::nextafter(x, INFINITY);
    return x;
}

hipError_t launch_solve(const SolveArgs& a_in, bool use6, bool def_f32, hipStream_t st, int phases) {
    SolveArgs a = a_in;
    a.stop2 = stop_threshold(a.tol);
    const int ppl = ppl_for(a.dim);
    if (ppl < 0) return hipErrorInvalidValue;
    if (phases & kSolveFast) hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, a.nbail, 1);
    if (def_f32 && a.defc && !use6) return launch_solve_cf32(a, ppl, st, phases);
    if (def_f32) return use6 ? launch_solve_kf32(a, ppl, st, phases) : launch_solve_bf32(a, ppl, st, phases);
    return use6 ? launch_solve_ku8(a, ppl, st, phases) : launch_solve_bu8(a, ppl, st, phases);
}

}  // namespace sdic
