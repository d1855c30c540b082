// CPU ORACLE — test infrastructure only. Nothing in the product links, loads or
// calls this code; tests, the smoke check and the bench's CPU-baseline leg use it
// as the checker.
//
// A plain C++17 restatement of the reference engine's full-field DIC solve
// (sempermechanics/semper-dic-engine 0.6.1), written from its sources:
//   src/math/image_processor.cpp        gradients and the three interpolators
//   src/math/subset_precomputer.cpp     pool entry (stats, Hessian, inverse), SD
//   src/math/hessian.hpp                rcond test
//   src/math/optimization_engine.cpp    IC-GN, Simplex, coarse search
//   include/semper/simd.hpp             the fused ZNSSD reductions
//   src/strain/strain_calculator.cpp    VSG strain and the median test
//   src/pipeline/full_field_*.cpp       validation, scan, seeding, paths A/B/C, pack
// Eigen and OpenCV are not available: the 6x6 / 3x3 inverses are written here
// (partial-pivot Gauss-Jordan, cofactors), and the OpenCV seeding front end
// (AKAZE homography + Delaunay) is replaced by the same global-motion search
// and lattice triangulation the GPU engine uses (their outputs are "parity
// unpinned" against the reference; everything downstream is pinned by the
// reference's golden fixtures).
// Parity unpinned, in particular: metrics slots 5-8 (Simplex calls and
// iterations per point), which depend on the anchors the replaced front end
// produces (FullFieldGolden: 1.66 vs 1.857 iterations per point, 8/7 vs 10/9
// Simplex calls on the affine scenario).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc {

// include/semper/tuning.hpp
constexpr float kCorrAccept = 0.15f;
constexpr float kCorrSimplexTrigger = 0.4f;
constexpr float kMinValidPixelFraction = 0.90f;
constexpr float kSimplexFailScore = 0.1f;
constexpr float kLmAlpha = 0.05f;
constexpr int kIcgnMaxIter = 50;
constexpr float kGhostRejectFraction = 0.05f;
constexpr float kAnchorSparseCoverage = 0.05f;
constexpr int kAnchorTarget = 256;
constexpr float kAnchorAcceptScore = 0.40f;
constexpr float kAnchorMedianTol = 1.0f;
constexpr float kAnchorFullFraction = 0.60f;
constexpr int kAnchorFullMinVertices = 10;
constexpr float kAnchorFullCoverage = 0.30f;
constexpr int kAnchorRefitMinPoints = 8;
constexpr float kSimilarityMinCornerPx = 4.0f;
constexpr float kSimilarityKeepFraction = 0.90f;
constexpr float kGhostWallValue = -10.0f;
constexpr float kGhostWallThreshold = -5.0f;
constexpr float kHessianMinRcond = 1e-6f;
constexpr float kCorrInvalid = -1.0f;
constexpr float kStrainUninit = -1000.0f;
constexpr float kStrainFail = -999.0f;
constexpr int kPathCMaxCandidates = 9;
constexpr int kPathBBatchNodes = 32;
constexpr double kVsgRadiusTiny = 1.0e-5;
constexpr int kOutlierRadius = 2;
constexpr float kOutlierEpsPx = 0.1f;
constexpr float kOutlierThreshold = 3.0f;
constexpr int kOutlierMinNeighbours = 3;
constexpr int kOutlierMaxPasses = 5;

constexpr int kErrRoi = -2, kErrInit = -3, kErrStrainWindow = -4, kErrSeed = -5,
              kErrCancelled = -99;

inline bool finite_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, sizeof u);
    return (u & 0x7f800000u) != 0x7f800000u;
}
inline bool finite_bits(double d) {
    uint64_t u;
    std::memcpy(&u, &d, sizeof u);
    return (u & 0x7ff0000000000000ull) != 0x7ff0000000000000ull;
}

// ---------------------------------------------------------------------------
// Image: float intensity plane plus 5-tap gradients (image_processor.cpp:9-64)
// ---------------------------------------------------------------------------
struct Image {
    int w = 0, h = 0;
    std::vector<float> I, gx, gy;
};
void image_from_u8(Image& img, int w, int h, const uint8_t* px);
void image_from_f32(Image& img, int w, int h, const float* px);
void image_gradients(Image& img);
float interp_bilinear(const Image& img, float x, float y);
float interp_bicubic(const Image& img, float x, float y);
float interp_keys6(const Image& img, float x, float y);

// ---------------------------------------------------------------------------
// 6x6 algebra (replacing Eigen): row-major float[36]
// ---------------------------------------------------------------------------
bool invert6(const float* A, float* Ainv);
bool invert6d(const double* A, double* Ainv);
bool well_conditioned(const float* H, int half);

// Pool entry: CachedHessianData (include/semper/subset.hpp:15-28)
struct PoolEntry {
    float H[36] = {};
    float Hinv[36] = {};
    float mean = 0.0f;
    float stdv = 1.0f;
    bool valid = false;
};
PoolEntry compute_hessian_only(const Image& ref, int cx, int cy, int dim);

// SubsetData (include/semper/types.hpp:60-100), SoA SD planes only
struct Subset {
    int cx = 0, cy = 0, dim = 0, n = 0;
    std::vector<float> ox, oy;       // float offsets, row-major -h..h
    std::vector<float> ref, nref;    // intensities and (I - mean)/std
    std::vector<float> gxv, gyv;     // raw reference gradients
    std::vector<float> sd;           // 6 planes of n
    float mean = 0.0f, stdv = 1.0f;
    float H[36] = {}, Hinv[36] = {};
    bool init = false;
};
void precompute_subset_fast(Subset& s, const Image& ref, int cx, int cy, int dim,
                            const PoolEntry& pe);
inline void precompute_subset(Subset& s, const Image& ref, int cx, int cy, int dim) {
    precompute_subset_fast(s, ref, cx, cy, dim, compute_hessian_only(ref, cx, cy, dim));
}

// AnalysisResult (include/semper/types.hpp:23-39)
struct Result {
    float u = 0, v = 0, ux = 0, uy = 0, vx = 0, vy = 0;
    int status = 0;
    float corr = 0.0f;
    int iters = 0;
    int invalid_ref = 0;
    int trigger_iters = 0;
};
enum Mode { AUTO_SEARCH = 0, NO_SEARCH = 1, NO_SIMPLEX = 2 };

// OptimizationEngine (include/semper/solver.hpp, src/math/optimization_engine.cpp)
struct Engine {
    bool lm = false;
    float lm_alpha = 0.0f;
    bool use6 = false;
    float conv_tol = 1e-3f;          // ||dp|| stop rule (optimization_engine.cpp:435)
    int simplex_calls = 0;
    double t_icgn_ms = 0.0, t_simplex_ms = 0.0;
    std::vector<float> buf_icgn, buf_simplex;
    std::vector<unsigned char> val_icgn, val_simplex;

    Result calculate_deformation(const Subset& s, const Image& def, const float g[6], Mode mode);
    Result solve_icgn(const Subset& s, const Image& def, const float g[6]);
    Result solve_simplex(const Subset& s, const Image& def, const Result& start, bool trans_only);
    float evaluate_znssd(const Subset& s, const Image& def, const float p[6]);
    void estimate_initial_guess(const Subset& s, const Image& def, float& bu, float& bv);
    float interp(const Image& def, float x, float y) const {
        return use6 ? interp_keys6(def, x, y) : interp_bicubic(def, x, y);
    }
};

// FP64 IC-GN of the same model: double image, double SD, double sums, iterated
// from `g` until ||dp|| < tol or max_iter. Returns status 0 on convergence.
struct Result64 {
    double p[6] = {};
    int status = 1;
    int iters = 0;
    double corr = 0.0;
};
Result64 icgn_fp64(const Image& ref, const Image& def, int cx, int cy, int dim, bool use6,
                   const double g[6], double tol, int max_iter);

// ---------------------------------------------------------------------------
// Strain (src/strain/strain_calculator.cpp)
// ---------------------------------------------------------------------------
struct DispField {
    int w = 0, h = 0, step = 0;
    std::vector<float> u, v;
    std::vector<unsigned char> valid;
};
struct StrainField {
    std::vector<float> exx, eyy, exy;
};
int vsg_window_node_count(int step, int window_px);
StrainField compute_vsg_strain(const DispField& d, int window_px);
int reject_displacement_outliers(DispField& d);
float median_inplace(float* v, size_t n);

// ---------------------------------------------------------------------------
// Full field (src/pipeline/full_field_*.cpp)
// ---------------------------------------------------------------------------
struct Params {
    int rect_x = 0, rect_y = 0, rect_w = 0, rect_h = 0;
    int step = 0, subset = 0, strain_window = 0;
    bool use6 = false;
};
int validate_params(Params& p, int width, int height);

struct RefCache {
    Image ref;
    int w = 0, h = 0;
    bool have = false;
    std::vector<uint8_t> gray;
    // derived-state cache (RefPrecompute): scan + pool
    bool grid_valid = false;
    Params grid_params;
    std::vector<uint8_t> grid_mask;   // mask the scan used (empty = none)
    int gridW = 0, gridH = 0;
    std::vector<unsigned char> skip;
    int total_valid = 0;
    std::vector<PoolEntry> pool;
};
void set_reference(RefCache& c, int w, int h, const uint8_t* gray, const uint8_t* mask);

// The solved node field of a run (FieldCapture, full_field_internal.hpp:247-268),
// in the form the incremental session samples: u .. corr of solved nodes with
// corr >= 0 (zero / kCorrInvalid elsewhere) and `valid` = those nodes after the
// displacement median test (node_field_from_capture, incremental_math.cpp:19-59).
struct Capture {
    int gw = 0, gh = 0, rect_x = 0, rect_y = 0, step = 0, attempted = 0;
    std::vector<float> u, v, ux, uy, vx, vy, corr;
    std::vector<unsigned char> valid;
};

struct RunOptions {
    int threads = 0;             // 0: hardware_concurrency
    float conv_tol = 1e-3f;
    Capture* capture = nullptr;  // filled when set
};
// Full-field solve. Returns point count or a negative code; metrics[24].
int run_full_field(RefCache& c, const uint8_t* def_gray, const uint8_t* mask, const Params& p,
                   float* out, int out_capacity, float* metrics, int metrics_len,
                   const RunOptions& opt);

// pack_full_field_output (full_field_solver_stats.cpp:21-91) of a node field
// on its own grid: the median test over `ok`, VSG strain with `window`, row-major
// pack with the capacity rule. Returns the point count.
int pack_field(const Capture& f, int window, float* out, int out_capacity);

// IncrementalSession (include/semper/incremental.hpp, src/pipeline/incremental_session.cpp).
struct IncInfo {
    int reference_frame = 0, reference_updated = 0, tracked_points = 0, frame0_points = 0,
        taylor_fallbacks = 0, lost_this_frame = 0;
};
struct IncSession;
IncSession* inc_create();
void inc_destroy(IncSession* s);
int inc_init(IncSession* s, const uint8_t* ref, int w, int h, const uint8_t* mask, const Params& p, int interval,
             const RunOptions& opt);
int inc_process(IncSession* s, const uint8_t* def, float* out, int out_capacity, float* metrics, int metrics_len,
                IncInfo* info);

}  // namespace orc
