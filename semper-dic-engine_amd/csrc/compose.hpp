// FP64 composition of an incremental session's totals on the device.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace sdic {

// The increment field of the last run (device): 7 planes of gw x gh, u v ux
// uy vx vy corr, and the flags the median test kept.
struct ComposeField {
    const float* planes = nullptr;
    const uint8_t* med = nullptr;
    int gw = 0, gh = 0, rect_x = 0, rect_y = 0, step = 0;
};

// The session's totals on its frame-0 grid (device): float and double, and
// the nodes still followed.
struct ComposeTotals {
    const float* uf = nullptr;
    const float* vf = nullptr;
    const double* u64 = nullptr;
    const double* v64 = nullptr;
    const uint8_t* alive = nullptr;
    int gw = 0, gh = 0, rect_x = 0, rect_y = 0, step = 0;
};

// Per node: flags bit 0 sampled, bit 1 by the first-order expansion; for a
// sampled node the new totals in double and rounded once (in float when the
// totals carry no double: u64 and v64 null in ComposeTotals and here), and
// the sample's corr (other nodes: flags 0, the rest untouched).
struct ComposeOut {
    double* u64 = nullptr;
    double* v64 = nullptr;
    float* uf = nullptr;
    float* vf = nullptr;
    float* corr = nullptr;
    uint8_t* flags = nullptr;
};

hipError_t launch_sample_compose64(const ComposeField& f, const ComposeTotals& t, const ComposeOut& o,
                                   hipStream_t st);

}  // namespace sdic
