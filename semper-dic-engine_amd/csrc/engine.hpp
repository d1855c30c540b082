// The MI355X full-field DIC engine behind the C ABI: one instance per
// semper_engine handle. It keeps the reference image, its gradient planes,
// the grid scan and the Hessian pool resident in HBM (the reference's
// ReferenceCache, include/semper/pipeline.hpp:35-64) and runs the solve of one
// deformed frame (run_full_field_impl, src/pipeline/full_field_solver.cpp:141-586)
// as a sequence of kernels with a thin host orchestration between them.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "compose.hpp"
#include "flood.hpp"
#include "seeder.hpp"
#include "solver.hpp"

namespace sdic {

// Device allocation owned by the engine; grows, never shrinks.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    bool reserve(size_t count) {
        if (count <= n && p) return true;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (hipMalloc(&p, sizeof(T) * (count ? count : 1)) != hipSuccess) {
            p = nullptr;
            return false;
        }
        n = count;
        return true;
    }
};

// Page-locked host memory (hipHostMalloc), for copies a captured graph
// replays: their host address is fixed in the graph.
template <typename T>
struct PinnedBuf {
    T* p = nullptr;
    size_t n = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    bool reserve(size_t count) {
        if (count <= n && p) return true;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
        if (hipHostMalloc((void**)&p, sizeof(T) * (count ? count : 1), hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            return false;
        }
        n = count;
        return true;
    }
};

struct Params {
    int rect_x = 0, rect_y = 0, rect_w = 0, rect_h = 0;
    int step = 0, subset = 0, strain_window = 0;
    bool use6 = false;
};

// The cancel token of a handle (include/semper/cancel.hpp:53-80).
struct Cancel {
    std::atomic<bool> flag{false};
    std::atomic<bool> window{false};
    void request() { flag.store(true, std::memory_order_relaxed); }
    bool requested() const { return flag.load(std::memory_order_relaxed); }
    void begin() {
        flag.store(false, std::memory_order_relaxed);
        window.store(true, std::memory_order_release);
    }
    void end() { window.store(false, std::memory_order_release); }
    void enter() {
        if (!window.exchange(false, std::memory_order_acq_rel)) flag.store(false, std::memory_order_relaxed);
    }
};

// Per-phase wall clock of the last solve (ms) and the device time of the
// fast-path solver launch over the grid, from HIP events on the engine's stream.
struct Timings {
    double total = 0, prep = 0, seeding = 0, prepass = 0, mesh = 0, path_a = 0, path_b = 0, strain = 0;
    double icgn = 0, simplex = 0;
    double kernel_path_a = 0;     // device ms of the Path A solver launch
    int path_a_items = 0;         // grid nodes the Path A launch covered
    int path_a_iters = 0;         // IC-GN iterations it ran
    int path_b_spec = 0;          // speculative first-claim solves of Path B (one launch)
    int path_b_launches = 0;      // further Path B solver launches (claims the speculation missed)
    int path_b_miss = 0;          // subsets those launches solved
};

// validate_params (full_field_solver.cpp:97-111) and
// vsg_window_node_count (strain_calculator.cpp:12-25).
int validate_params(Params& p, int width, int height);
int vsg_window_node_count(int step, int window);

class Engine {
public:
    Engine();
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    void reset();
    bool has_reference() const { return have_ref_; }
    int width() const { return w_; }
    int height() const { return h_; }

    // ReferenceCache::set_from_gray: u8 reference (host or device) and an
    // optional ghost-wall mask (host). 0 or a negative code.
    int set_reference(const uint8_t* gray, bool on_device, int w, int h, const uint8_t* mask);

    // run_full_field on a u8 frame of the reference's size (host or device);
    // the scan mask (host, may be null). Returns the point count or a code.
    int run(const uint8_t* def, bool on_device, const uint8_t* mask, const Params& p, float* out,
            int out_capacity, float* metrics, int metrics_len, Cancel& cancel,
            const std::function<void(int)>& progress, float conv_tol = 1e-3f);

    const Timings& timings() const { return tm_; }
    hipStream_t stream() const { return st_; }

    // Device planes of the last solve (for incremental sessions): per grid node.
    struct Field {
        int gw = 0, gh = 0, rect_x = 0, rect_y = 0, step = 0, attempted = 0;
        std::vector<float> u, v, ux, uy, vx, vy, corr;
        std::vector<unsigned char> ok;   // solved with corr >= 0 (before the median test)
        std::vector<unsigned char> med;  // ok and kept by the displacement median test
    };
    // A stored subset to match on the last run's deformed frame
    // (match_anchor, src/pipeline/incremental_math.cpp:237-272): a gray patch
    // of side 2 (subset / 2 + margin) + 1 centred on (ax, ay) of the frame it
    // was cut from, the node at (ax + dx, ay + dy) there, and the guess of
    // where the node is now.
    struct AnchorMatch {
        int slot = -1;            // the patch's slot in the caller's device store
        int side = 0, ax = 0, ay = 0;
        float dx = 0, dy = 0, guess_x = 0, guess_y = 0, max_shift = 0;
        bool ok = false;          // out: converged, ZNSSD <= 0.15, within max_shift
        float x = 0, y = 0, corr = 0;
    };
    // Matches every request with the subset solver (NO_SEARCH with the Simplex
    // rescue, LM damping) on the deformed frame of the last run; the patches
    // are slots of `store` (side x side bytes each, device memory). Returns 0
    // or a negative code.
    int match_anchors(std::vector<AnchorMatch>& reqs, const uint8_t* store, int subset, bool use6, float tol);
    // Copies side x side blocks of the last run's deformed frame into slots of
    // a device store: 3 ints per block (slot, left column, top row), each
    // block inside the frame. Enqueued on the engine's stream. 0 or kErrInit.
    int cut_patches(const std::vector<int32_t>& entries, int side, uint8_t* store);
    bool keep_field = false;
    // Solve Path B's first claims speculatively in one launch (off: every round
    // is solved when it is reached; the committed field is the same either way).
    bool speculate_path_b = true;
    // The layered Path B's frontier (sources, search, tasks, updates) on the
    // device (on by default, for reach <= 16 on one subset size; off: on the
    // host; the field is the same either way).
    bool flood_on_device = true;
    // Layered Path B (no reference counterpart; the default in learned
    // seeding, with reach 8): instead of the reference's priority flood, which
    // claims up to 32 nodes per round in corr order, each layer claims at once
    // every open node within `reach` grid steps of an accepted one, from the
    // accepted node a breadth-first search in (corr, y, x) order reaches it
    // from first, carrying that node's warp over the offset, and solves them
    // all in one launch. Adjacent claims solve as the flood's do (Simplex
    // rescue on, except under learned seeding) and a failure there is final;
    // farther claims run without
    // the rescue and a failure leaves the node open for an adjacent claim. So
    // a node is claimed at most twice, and the launches number about the
    // depth of the unsolved regions over `reach`. 0: the reference flood.
    int layered_path_b = 0;
    // learned seeding's reach: on the 4096^2 bumps frame reach 8 took 12
    // layers and 6.0 ms of Path B, 16 takes 7 layers and 5.3 ms, with the
    // same points and accuracy
    static constexpr int kLayeredReachLearned = 16, kLayeredReachMax = 32;
    // `out` of run() is device memory (the packed points stay in HBM)
    bool out_on_device = false;
    // Precision variants (no reference counterpart): the Gauss-Newton step by
    // FP64 Cholesky of each pool entry, fp16 storage of the reference gradient
    // planes the pool and the solver read, and J^T J on MFMA in the pool.
    // A change rebuilds the grid cache on the next run.
    struct Precision {
        bool chol64 = false;
        bool grad16 = false;
        bool grad_bf16 = false;   // with grad16: bf16 storage instead of fp16
        bool mfma_pool = false;
        bool operator==(const Precision& o) const {
            return chol64 == o.chol64 && grad16 == o.grad16 && grad_bf16 == o.grad_bf16 && mfma_pool == o.mfma_pool;
        }
    };
    Precision prec;
    // Learned anchor seeding (no reference counterpart): with a seeder loaded,
    // each lattice anchor's guess comes from the CNN (relative to the global
    // model's shift), IC-GN refines it (NO_SEARCH: the Simplex rescue when it
    // fails), and an anchor whose refined solve fails is solved again the
    // plain way.
    bool learned_seed = false;
    // Learned seeding runs the CNN lattice on every frame (off: only where
    // the global model finds a large motion, or the plain lattice's anchors
    // do not converge).
    bool cnn_always = false;
    // The seeder's inference: the fp16 MFMA kernel (true) or the FP32
    // forward pass it is trained with.
    bool seeder_mfma = true;
    // A CNN-seeded anchor whose IC-GN fails takes the Simplex rescue before
    // the plain solve (true), or goes to the plain solve at once (false:
    // measured slower, as the anchors the rescue would hold are lost and the
    // lattice misses its quality gate and is solved a second time).
    bool cnn_rescue = true;
    // The solver samples an FP32 copy of the deformed frame (made once per
    // run; the values are the u8 ones, so the field is the same bit for bit)
    // instead of unpacking four bytes per row of every Keys gather.
    bool def_f32 = true;
    // Keys 4x4 from a row-cubic coefficient plane of the deformed frame (no
    // reference counterpart; off by default): made once per run (per pixel the
    // cubics of two rows, four fp16 each, exact for u8 frames), read by the
    // fast path's interior samples with two 16-byte loads per sample and a
    // Horner cubic per stencil row, instead of four row loads and four taps
    // weighted (half the vector-memory instructions, a quarter fewer
    // vector-ALU ones). Needs def_f32 (the full solver keeps the FP32 plane)
    // and a frame of under 2^27 pixels (32-bit byte offsets).
    bool def_coef = false;
    // The global-motion search that seeds the anchor lattice (on by
    // default). Off, the lattice and Path C start from a zero shift, as the
    // reference's do when its feature match finds no model.
    bool global_search = true;
    // Mixed subset sizes (no reference counterpart): the subset size of each
    // node of a gw x gh grid (row-major), used by runs on a grid of that shape
    // (any other shape, or an empty map: the run's subset everywhere). Each
    // size odd, 3 .. 63. The reference's semantics apply node by node with
    // the node's own size. 0 or -3.
    int set_subset_map(int gw, int gh, const std::vector<int>& dims);
    // The same by image region: an rw x rh map (row-major) over the frame; a
    // node at pixel (x, y) takes the size of region (x rw / w, y rh / h), for
    // any grid, so it follows grids whose rect changes (the incremental
    // session's). A per-node map of the run's grid shape takes precedence. An
    // empty map clears it. 0 or -3.
    int set_region_map(int rw, int rh, const std::vector<int>& dims);
    // A size map's shape and sizes are acceptable (an empty map is).
    static bool size_map_ok(int cols, int rows, const std::vector<int>& dims);
    // Launch-shape tuner (no reference counterpart). A shape -- 1, 2 or 4
    // subsets per fast-path workgroup, row-major or 8 x 8-tile node order --
    // changes which wave solves which subset, never a subset's arithmetic, so
    // the first runs of each (grid, subset size, interpolator, item count)
    // rotate through the shapes as ordinary launches, kTuneSamples each, timed
    // in place by device events; from then on the shape with the least mean
    // time is used.
    static constexpr int kTuneSamples = 2;
    bool autotune = true;
    int force_shape = 0;   // 1..6: use that candidate everywhere (tests), 0: tune
    struct Tune {
        int wpb = 1, tile = 0;
    };
    struct TuneReport {
        bool have = false, settled = false;
        Tune chosen;
        double ms[6] = {0, 0, 0, 0, 0, 0};   // mean per candidate, -1 when not (yet) timed
    };
    TuneReport last_tune;
    // Graph capture of a frame's fixed kernel chains (no reference
    // counterpart): the end of a run (valid flags, the five median passes,
    // VSG, the row-major pack and the count download) and the session's
    // strain of the total field (uploads, median passes, VSG, downloads). A
    // chain seen twice in a row with the same grid, parameters and buffers
    // is captured on the second call and replayed from then on; host
    // transfers go through page-locked staging, counters are cleared by a
    // kernel, never a memset.
    bool use_graphs = true;
    // Per-node provenance of the last run (the reference's debug export,
    // full_field_debug_export.cpp, in a form of its own): for each grid node
    // the mesh tag of its guess (Tag) and the path that solved it (NodePath).
    // Costs two grid-sized copies per run while on.
    bool node_report = false;
    enum NodePath : uint8_t { kPathSkipped = 0, kPathUnsolved = 1, kPathAnchor = 2, kPathA = 3, kPathLater = 4 };
    std::vector<uint8_t> report_tag, report_path;
    // with node_report: the seven node planes (u, v, ux, uy, vx, vy, corr,
    // plane after plane) after Path A and at the end of the run
    std::vector<float> report_planes_a, report_planes;
    // and the six guess planes (u, v, ux, uy, vx, vy) Path A started from
    std::vector<float> report_guess;
    // The global-motion model of the last run: locked, full, a0..a5, cx, cy,
    // patches, the pass that locked it (0 none, 1 translation search, 2 the
    // wide translation search, 3 the rotation-invariant search).
    double last_model[12] = {};
    struct GraphStats {
        int captures = 0, replays = 0;
    };
    GraphStats graph_stats() const;
    int load_seeder(const std::string& path);
    bool seeder_ready() const { return seeder_ && seeder_->ready(); }
    struct SeedStats {
        int cnn = 0, cnn_ok = 0, fallback = 0;
    };
    SeedStats seed_stats;
    // (work items, items the fast path handed to the full solver) of each
    // subset-solver launch of the last run, in launch order; a Path A launch
    // over several subset sizes reports the hand-offs of all of them
    std::vector<std::pair<int, int>> launch_log;
    Field last_field;
    // The same field on the device when keep_field (the node planes and the
    // median test's flags; valid until the next run or reference change).
    ComposeField last_dev_field;

    // Median test, VSG strain and row-major packing of a node field given on
    // the host (f.ok marks the measured nodes; corr is reported as is). The
    // incremental session's output path. Returns the point count.
    int strain_pack(const Field& f, int strain_window, float* out, int out_capacity, int* dropped_post,
                    int* dropped_outlier);

private:
    int ensure_grid(const Params& p, const uint8_t* mask, const std::vector<int>* dims, bool& cached);
    int solve_tasks(const std::vector<Task>& tasks, std::vector<SolveOut>& out, bool use6, int dim, float tol);

    hipStream_t st_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    bool have_ref_ = false;
    int w_ = 0, h_ = 0;
    DevBuf<uint8_t> ref_u8_, def_u8_;
    DevBuf<float> def_f32_;      // the deformed frame as FP32, when def_f32
    bool def_f32_ready_ = false;
    const void* def_plane() const { return def_f32_ready_ ? (const void*)def_f32_.p : (const void*)def_u8_.p; }
    DevBuf<uint4> def_coef_;     // row-cubic coefficient plane of the frame, when def_coef
    DevBuf<int32_t> stat_slots_; // Path A's counter slots (kStatSlots x kStatStride)
    bool def_coef_ready_ = false;
    void set_def_args(SolveArgs& a) const {
        a.def = def_plane();
        a.defc = def_coef_ready_ ? def_coef_.p : nullptr;
    }
    DevBuf<float> I_, gx_, gy_;
    DevBuf<float> pyr_r_[4], pyr_d_[4];
    int pyr_ref_levels_ = 0;
    // grid cache (RefPrecompute)
    bool grid_valid_ = false;
    Params grid_p_;
    std::vector<uint8_t> grid_mask_;
    int gw_ = 0, gh_ = 0, total_valid_ = 0;
    std::vector<uint8_t> skip_h_;
    DevBuf<uint8_t> skip_d_, mask_d_;
    DevBuf<PoolEntry> pool_;
    DevBuf<uint8_t> pool_raw_;      // per-node sums between the two pool launches
    Precision grid_prec_;
    std::unique_ptr<Seeder> seeder_;
    std::vector<int> subset_map_;
    int map_gw_ = 0, map_gh_ = 0;
    std::vector<int> region_map_, region_dims_;
    int region_w_ = 0, region_h_ = 0;
    std::vector<int> grid_dims_;        // the grid cache's per-node sizes (empty: uniform)
    DevBuf<int32_t> dims_d_;
    struct DimClass {
        int dim, off, n;
    };
    std::vector<DimClass> classes_;     // Path A launches: nodes of one size each
    DevBuf<int32_t> class_nodes_d_;
    int solve_tasks_dim(const std::vector<Task>& tasks, std::vector<SolveOut>& out, bool use6, int dim, float tol);
    // Two task sets over the same n items, the second needed only where the
    // first fails (one subset size): the first set's fast path, then its full
    // solver on the engine's stream while the second set solves, on a stream
    // of its own, the items that fast path handed on -- every item the full
    // solver can still fail. o2[k] is set where handed[k]; each result is the
    // one a launch of that task alone gives.
    int solve_tasks_spec(const std::vector<Task>& t1, const std::vector<Task>& t2, std::vector<SolveOut>& o1,
                         std::vector<SolveOut>& o2, std::vector<uint8_t>& handed, bool use6, int dim, float tol);
    struct TuneState {
        int next = 0;                      // runs so far
        int chosen = -1;                   // settled candidate
        std::vector<float> samples[6];
    };
    std::map<std::vector<int>, TuneState> tuned_;
    std::vector<hipEvent_t> class_ev_;     // before and after each class launch, then the fork and the join
    std::vector<hipStream_t> class_st_;    // the streams of the Path A classes after the first
    DevBuf<int32_t> class_nbail_;          // their hand-off counters
    // events, streams and counters for launches over nc subset sizes at once
    void ensure_class_streams(size_t nc);
    // The candidate for the next Path A launch of a size class (index into the
    // candidate list), and the report of a finished timing.
    int shape_for(const std::vector<int>& key, bool listed);
    void shape_timed(const std::vector<int>& key, bool listed, int cand, float ms);
    DevBuf<uint16_t> gx16_, gy16_;  // fp16 gradient planes (Precision::grad16)
    int g16_fmt_ = -1;   // format of the 16-bit gradient planes made (-1 none, 0 fp16, 1 bf16)
    DevBuf<double> chol_;           // FP64 Cholesky factors (Precision::chol64)
    // the precision-variant pointers of a solver launch on the grid pool
    void set_precision_args(SolveArgs& a) const;
    // per run
    DevBuf<float> gplanes_;        // u v ux uy vx vy corr  (7 planes)
    DevBuf<uint8_t> gflags_;       // solved tag simplex in_mesh valid ok_a ok_b (7 planes)
    DevBuf<int32_t> giters_;
    DevBuf<float> guess_, tmp_;    // 6 planes each
    DevBuf<float> strain_;         // exx eyy exy
    struct Chain {
        std::vector<uint64_t> key;
        hipGraphExec_t exec = nullptr;
        int captures = 0, replays = 0;
    };
    Chain tail_chain_, strain_chain_;
    std::vector<uint8_t> after_a_;   // solved flags after Path A, when node_report
    PinnedBuf<int32_t> pin_counts_;   // the counters after a run's chain
    PinnedBuf<float> pin_in_, pin_out_;
    PinnedBuf<uint8_t> pin_ok_in_, pin_ok_out_;
    PinnedBuf<float> pin_planes_;    // Path B's copy of the node planes
    PinnedBuf<uint8_t> pin_flags_;
    struct FloodScratch {   // the layered Path B's per-node state
        std::vector<uint8_t> ok, fin;
        std::vector<int8_t> tried;
        std::vector<int> stamp, dist, src, commit_at;
        int epoch = 0;   // stamps below it belong to earlier layers or runs
    };
    FloodScratch flood_;
    // The device frontier of the layered Path B (flood.hpp): per-node flags,
    // warps, search state and results, the sort's scratch, the counters.
    DevBuf<uint8_t> fl_u8_;            // ok, fin, committed (3 NN), tried (NN)
    DevBuf<float> fl_w_;               // 7 NN
    DevBuf<int> fl_int_;               // stamp dist cand sources level_a level_b reached tdist commit_nodes (9 NN)
    DevBuf<unsigned long long> fl_key_;   // key, skeys, skeys_sorted (3 NN)
    DevBuf<Task> fl_tasks_;
    DevBuf<SolveOut> fl_res_;          // a layer's results, per-node results, the commit list (3 NN)
    DevBuf<int> fl_counters_;
    DevBuf<uint8_t> fl_sort_temp_;
    PinnedBuf<int32_t> pin_fl_;        // the counters and the hand-off count read back
    FloodDev flood_dev(int gw, int gh);
    // The subset solver over tasks already on the device (no host copies):
    // enqueued on the engine's stream; the hand-off count lands in
    // counters_[8].
    int solve_tasks_device(const Task* tasks, int n, SolveOut* out, bool use6, int dim, float tol);
    PinnedBuf<Task> pin_tasks_;        // a task launch's upload
    PinnedBuf<SolveOut> pin_res_;      // and its results
    PinnedBuf<int32_t> pin_handed_;    // and its hand-off count
    template <typename F>
    void run_chain(Chain& c, const std::vector<uint64_t>& key, F&& enqueue);
    DevBuf<float> pack_out_;       // packed output points, 8 floats each
    DevBuf<int32_t> pack_blocks_;  // per-block write counts of the pack
    DevBuf<Task> tasks_d_;
    DevBuf<int32_t> patch_list_;    // cut_patches entries, match_anchors slots
    DevBuf<SolveOut> out_d_;
    DevBuf<int32_t> bail_, counters_, nodes_d_;
    DevBuf<TriDev> tris_d_;
    DevBuf<int32_t> celltri_d_, lat_d_;
    DevBuf<int32_t> centres_d_;
    DevBuf<PatchOut> patch_d_;
    DevBuf<uint8_t> patch_scratch_;   // the coarsest level's shared candidates
    // the rotation-invariant pass: level centres, descriptions and
    // orientations, scores, best matches, the refinement's start shifts
    DevBuf<int32_t> canon_centres_d_, canon_start_;
    DevBuf<float> canon_desc_, canon_scores_;
    DevBuf<CanonOut> canon_out_;
    Timings tm_;
};

}  // namespace sdic
