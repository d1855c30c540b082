// Subset-solver instantiations: bicubic 4x4 interpolation over a uint8_t deformed plane.
#include "solver_impl.hpp"

namespace sdic {

hipError_t launch_solve_bu8(const SolveArgs& a, int ppl, hipStream_t st, int phases) {
    if (a.nitems <= 0) return hipSuccess;
    const int wpb = a.wpb == 2 || a.wpb == 4 ? a.wpb : 1;
    const dim3 block(64 * kSlowWaves), fast_block(64 * wpb), fast_grid((a.nitems + wpb - 1) / wpb),
        slow_grid(a.nitems < 2048 ? a.nitems : 2048);
    switch (ppl) {
#define SDIC_CASE(P)                                                          \
    case P:                                                                   \
        if (phases & 1) hipLaunchKernelGGL((k_fast<P, false, uint8_t>), fast_grid, fast_block, 0, st, a); \
        if (phases & 2) hipLaunchKernelGGL((k_slow<P, false, uint8_t, kSlowWaves>), slow_grid, block, 0, st, a); \
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

}  // namespace sdic
