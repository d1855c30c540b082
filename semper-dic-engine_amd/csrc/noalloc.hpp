// The subset solver without allocation (csrc/noalloc.hip): its limits and the
// host entry behind sdk_solve_nodes_noalloc.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace sdic {

// Largest image (w * h) and node count one call takes; the fixed device array
// holds three reference planes of kNoAllocMaxPixels floats and, per node, a
// pool entry, a task, a result and a hand-off slot.
constexpr size_t kNoAllocMaxPixels = (size_t)1 << 20;
constexpr int kNoAllocMaxNodes = 16384;
constexpr size_t kNoAllocScratchBytes = 17105152;   // 16.3 MiB
constexpr int kNoAllocErrArgs = -3;
constexpr int kNoAllocErrCapacity = -7;

// Enqueues the solve on st; 0, kNoAllocErrArgs or kNoAllocErrCapacity (nothing
// enqueued). Every pointer is device memory.
int solve_nodes_noalloc(const float* ref, const void* def, int def_is_f32, int w, int h, int dim, int use6, int lm,
                        float tol, int n, const int32_t* cxy, const float* guess, const int32_t* modes,
                        int32_t* iout, float* fout, hipStream_t st);

}  // namespace sdic
