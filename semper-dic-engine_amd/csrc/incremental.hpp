// Incremental (updated-reference) DIC over a frame sequence
// (include/semper/incremental.hpp:74-117, src/pipeline/incremental_session.cpp).
//
// Each frame is solved on the GPU against the current reference on that
// reference's grid; the increment is composed onto the running total at the
// frame-0 grid points in FP64,
//     U(X0, t) = U(X0, k) + u_k(X0 + U(X0, k)),
// with u_k sampled bilinearly between solved nodes, or by the first-order
// expansion of the nearest solved node within 1.5 cells near an edge. Every
// `update_interval` frames (0: never) the frame just solved becomes the
// reference; a frame-0 point that could not be sampled at an update is lost.
// Output: the frozen 8-float layout at the frame-0 grid, Green-Lagrange strain
// of the total displacement.
//
// Covered: the composition, the Taylor fallback, the update interval and the
// update floor (no move to a frame reporting under half the points of the
// reference's first frame). Not covered yet: coasting of unsampled points
// across frames, re-anchoring, the early update on a 5% drop, and the mask warp.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <vector>

#include "engine.hpp"

namespace sdic {

struct FrameInfo {
    int reference_frame = 0;
    bool reference_updated = false;
    int tracked_points = 0;
    int frame0_points = 0;
    int taylor_fallbacks = 0;
    int lost_this_frame = 0;
};

class IncrementalSession {
public:
    void reset();
    bool initialized() const { return init_; }
    int init(const uint8_t* ref, int w, int h, const uint8_t* mask, const Params& p, int update_interval);
    int process(const uint8_t* def, float* out, int out_capacity, float* metrics, int metrics_len, Cancel& cancel,
                const std::function<void(int)>& progress, FrameInfo* info);

private:
    std::unique_ptr<Engine> eng_;
    bool init_ = false;
    int w_ = 0, h_ = 0, interval_ = 0, frame_ = 0, ref_frame_ = 0;
    int ref_first_count_ = -1;     // points reported by the first frame on the current reference
    int ref_own_count_ = 0;        // points the reference reported when it was solved (frame 0: scan count)
    Params p_;
    std::vector<uint8_t> mask0_;
    int gw_ = 0, gh_ = 0, frame0_points_ = 0;
    std::vector<uint8_t> alive_;
    std::vector<double> ub_, vb_;  // total displacement of each frame-0 node at the current reference
};

}  // namespace sdic
