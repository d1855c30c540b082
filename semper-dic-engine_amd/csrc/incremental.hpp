// Incremental (updated-reference) DIC over a frame sequence
// (include/semper/incremental.hpp:74-117, src/pipeline/incremental_session.cpp,
// src/pipeline/incremental_math.cpp).
//
// Every frame is solved by the full-field engine on the GPU against the
// current reference, on that reference's own grid; each frame-0 node's
// increment is sampled where the node sits on the reference and added to its
// total,
//     U(X0, t) = U(X0, k) + u_k(X0 + U(X0, k)),
// bilinearly between solved nodes, or by the first-order expansion of the
// nearest solved node within 1.5 cells. A node the solve does not reach coasts,
// unreported, on a weighted plane through the solved increments within 8 cells
// (at most 2 subsets of travel); a node that coasted further than a quarter
// subset is reported again only once its subset, stored where it was last
// clean, matches the frame. The reference moves to the frame just solved every
// `update_interval` frames (0: never), early when the points reported fall 5%
// below the first frame on the current reference, and never to a frame
// reporting under half of that count (or of the reference's own, or of the
// frame-0 scan). At an update the frame-0 material mask is warped forward into
// the new reference and the new grid spans where the tracked nodes sit. Output:
// the frozen 8-float layout at the frame-0 grid, Green-Lagrange strain of the
// total displacement, corr of the increment. While the reference is frame 0,
// a frame is exactly the engine's run on it.
#pragma once

#include <cstdint>
#include <functional>
#include <map>
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
    int frames_processed() const { return frame_count_; }
    int init(const uint8_t* ref, int w, int h, const uint8_t* mask, const Params& p, int update_interval);
    int process(const uint8_t* def, float* out, int out_capacity, float* metrics, int metrics_len, Cancel& cancel,
                const std::function<void(int)>& progress, FrameInfo* info);

    // The solver's stop rule (the reference's fixed 1e-3 through the C ABI).
    float tol = 1e-3f;
    // Mixed subset sizes by image region (Engine::set_region_map; no
    // reference counterpart): every solve of the session, on the first
    // reference and on each adopted one, takes each node's size from the
    // region of the frame it lies in. Kept across init. 0 or -3.
    int set_region_map(int rw, int rh, const std::vector<int>& dims);
    // Graph capture of the engine's fixed per-frame chains (Engine::use_graphs;
    // kept across init), and the captures and replays of the current engine.
    bool use_graphs = true;
    // FP64 composition of the total displacement (no reference counterpart;
    // the reference adds each increment to a float total): the totals carried
    // across reference updates are kept in double, each increment is added in
    // double, and the reported total is that sum rounded once. Kept across init.
    bool compose64 = false;
    // Host milliseconds of the last process call: total, the engine's run,
    // the composition of the totals (sampling, re-anchoring, strain of the
    // total field), the reference update, and within it the field fill, the
    // mask warp, the warped mask's bounds and the adoption of the frame.
    double timing[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // Stored subsets the last process call matched on its frame (nodes that
    // coasted past a quarter subset and came back), and of those the ones
    // accepted.
    int reanchor_requests = 0, reanchor_matched = 0;
    Engine::GraphStats graph_stats() const { return eng_ ? eng_->graph_stats() : Engine::GraphStats{}; }
    // the session's engine (null before init)
    const Engine* engine() const { return eng_.get(); }

private:
    std::vector<int> region_map_;
    int region_w_ = 0, region_h_ = 0;
    // A stored subset (make_anchor, incremental_math.cpp:224-235): its gray
    // patch of side anchor_side_ around (ax, ay) lives in slot `slot` of the
    // device store patch_store_ (-1: the node has none), cut from the frame
    // that became the reference by one launch per update.
    struct Anchor {
        int slot = -1;
        int ax = 0, ay = 0;
        float dx = 0, dy = 0;
    };

    std::unique_ptr<Engine> eng_;
    bool init_ = false;
    int w_ = 0, h_ = 0, interval_ = 0;
    Params p0_;                       // the frame-0 grid (validated)
    int gw_ = 0, gh_ = 0;
    std::vector<uint8_t> user_mask0_, material_mask0_;
    int frame0_points_ = -1;
    std::vector<float> ub_, vb_, coast_travel_;
    std::vector<double> ub64_, vb64_;   // the same totals in double, when c64_
    bool c64_ = false;                   // compose64 as it was at init
    // The device side of the FP64 composition: the totals in (float, then
    // double) and out, the followed nodes, the per-node sample flags and corr.
    DevBuf<float> c64_uf_;               // 4 n0: uf vf in, uf vf out
    DevBuf<double> c64_u64_;             // 4 n0: u64 v64 in, u64 v64 out
    DevBuf<float> c64_corr_;
    DevBuf<uint8_t> c64_flags_;          // 2 n0: alive in, flags out
    // Samples every followed node's increment and adds it to its total in
    // double on the device (launch_sample_compose64); host copies of the
    // results for the node loop.
    int compose_on_device(std::vector<uint8_t>& flags, std::vector<double>& u64, std::vector<double>& v64,
                          std::vector<float>& uf, std::vector<float>& vf, std::vector<float>& corr);
    std::vector<uint8_t> alive_;
    std::vector<Anchor> anchors_;         // per frame-0 node
    std::vector<int32_t> free_slots_;     // store slots released by dropped anchors
    int slots_used_ = 0, anchor_side_ = 0;
    DevBuf<uint8_t> patch_store_;         // slots x side x side bytes
    void drop_anchors();
    int ref_index_ = 0, frames_since_update_ = 0, ref_reported_ = -1, adopted_reported_ = -1, frame_count_ = 0;
    Params pk_;                       // the grid on the current reference (k > 0)
    std::vector<uint8_t> mask_k_;
    std::vector<float> scratch_;

    int alive_count() const;
    int nothing_to_follow(FrameInfo fi, float* metrics, int metrics_len, const std::function<void(int)>& progress,
                          FrameInfo* info);
};

}  // namespace sdic
