// Shared host/device definitions of the MI355X DIC engine: the solver constants
// of the reference engine (include/semper/tuning.hpp), and the plain-old-data
// records that live in HBM and cross between host code and kernels.
#pragma once

#include <cstdint>

namespace sdic {

// include/semper/tuning.hpp (values are the reference's)
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

// Largest subset the register-resident solver takes: 63 x 63 = 3969 pixels,
// 64 per lane of a 64-wide wavefront.
constexpr int kMaxSubset = 63;

enum Mode : int32_t { AUTO_SEARCH = 0, NO_SEARCH = 1, NO_SIMPLEX = 2 };

// How a node got its guess (full_field_internal.hpp:177-183).
enum Tag : int32_t { kMeshNone = 0, kMeshInTriangle = 1, kMeshExtrapolated = 2, kMeshAnchor = 3,
                     kMeshHomography = 4 };

// One grid node's Hessian pool entry (CachedHessianData): the step matrix the
// IC-GN uses on fully valid iterations -- H^-1, or (H + diag(a, a, 0..))^-1 when
// the pool was built for the LM-damped full-field engine -- plus the subset
// statistics. 160 bytes, read with scalar loads by the wave that solves the node.
struct alignas(16) PoolEntry {
    float hstep[36];
    float mean;
    float inv_std;
    int32_t valid;
    int32_t pad;
};
static_assert(sizeof(PoolEntry) == 160, "pool entry layout");
// Doubles per pool entry of the optional FP64 Cholesky factors (21 used: the
// lower triangle, its diagonal as reciprocals).
constexpr int kCholStride = 24;

// One solve request: node centre, its pool slot, a 6-DOF guess and the
// initialisation mode (calculate_deformation's arguments).
struct alignas(16) Task {
    int32_t cx, cy;
    int32_t pool;
    int32_t mode;
    float g[6];
    int32_t node;   // grid index the result belongs to (-1: none)
    int32_t pad;
};
static_assert(sizeof(Task) == 48, "task layout");

// AnalysisResult (include/semper/types.hpp:23-39) plus the solve-site flags.
struct alignas(16) SolveOut {
    float p[6];          // u, v, ux, uy, vx, vy
    float corr;
    int32_t status;
    int32_t iters;
    int32_t trigger_iters;
    int32_t simplex;     // a Simplex ran
    int32_t init;        // the subset and its pool entry were valid
    int32_t invalid_ref;
    int32_t pad[3];
};
static_assert(sizeof(SolveOut) == 64, "result layout");

// Per-node state of the solve on the device (the ResultGrid as planes).
struct GridDev {
    float* u;
    float* v;
    float* ux;
    float* uy;
    float* vx;
    float* vy;
    float* corr;
    uint8_t* solved;
    uint8_t* tag;
    int32_t* iters;
    uint8_t* simplex;   // bit0: a Simplex ran; bit1: it was recorded as saved
};

// Lattice triangle of the seeding mesh: vertex positions in reference pixels,
// the affine guess it carries, and its bounding box padded by 15 px.
struct alignas(16) TriDev {
    float px[3], py[3];
    double ux, uy, vx, vy, u, v;
    float bx0, by0, bx1, by1;
    int32_t cell;
    int32_t pad[3];
};

}  // namespace sdic
