// Forward warp of a material mask through a node-grid displacement field
// (incremental session reference updates); see maskwarp.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace sdic {

struct MaskGrid {
    int gw, gh, rx, ry, step;
};

// mask0: W x H host bytes (>= 128 = material). U, V: the filled grid field.
// [X0, X1] x [Y0, Y1]: the frame-0 window. Writes the closed, warped mask
// (255 / 0) to out. Returns 0 or -3.
int mask_warp(const uint8_t* mask0, int W, int H, const MaskGrid& g, const std::vector<float>& U,
              const std::vector<float>& V, int X0, int Y0, int X1, int Y1, int per_px, std::vector<uint8_t>& out,
              hipStream_t st);

}  // namespace sdic
