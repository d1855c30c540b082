// Launch interface of the device kernels (host side). All pointers are device
// pointers; every launcher is asynchronous on the given stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace sdic {

// Path A over the grid: node -> centre, guess planes, tag.
struct GridArgs {
    GridDev grid;
    const uint8_t* in_mesh;
    const float *gu, *gv, *gux, *guy, *gvx, *gvy;
    int gw, gh, rect_x, rect_y, step, dim, w, h;
    int skip_model;   // leave nodes guessed from the global model to a later path
    int no_rescue;    // no Simplex rescue: a failure is left to a later path
};

// Optional precision variants of the Hessian pool and the solver: J^T J on
// MFMA (fp16 steepest-descent values, FP32 accumulation), fp16 storage of the
// reference gradient planes (half bits in uint16_t), and the FP64 Cholesky
// factor of each pool entry for the Gauss-Newton step.
struct PoolOpts {
    bool mfma = false;
    const uint16_t* gx16 = nullptr;
    const uint16_t* gy16 = nullptr;
    bool bf16 = false;   // the 16-bit gradient copies are bf16 (else fp16)
    double* chol = nullptr;
    const int* dims = nullptr;   // per-node subset sizes (mixed subsets; else `dim`)
};

// Arguments of the subset solver kernel. Task mode when tasks != null (one
// result per task); grid mode (Path A) otherwise.
struct SolveArgs {
    const Task* tasks;        // task mode: work item id indexes tasks[] and out[]
    SolveOut* out;
    const int* items;         // optional list of work item ids (null: 0 .. nitems-1)
    int nitems;
    int* bail;                // ids the fast path hands to the full solver
    int* nbail;               // their count (device)
    GridArgs grid;
    int* stats;               // grid mode: kStatSlots slots of kStatStride ints (points, iterations,
                              // simplex calls, saved, dead), a workgroup counting into slot
                              // blockIdx.x % kStatSlots; launch_sum_stats adds them up
    const PoolEntry* pool;
    const float* refI;
    const float* refGx;
    const float* refGy;
    const uint16_t* refGx16;  // 16-bit gradient planes (read instead of refGx / refGy when set)
    const uint16_t* refGy16;
    int grad_bf16;            // those planes hold bf16 (else fp16)
    const double* chol;       // FP64 Cholesky factors per pool entry (the step when set)
    const void* def;          // uint8 or float plane
    const uint4* defc;        // row-cubic coefficient plane of the same frame (Keys 4x4 fast path), or null
    int w, h, dim, lm;
    float tol;
    // set by launch_solve: the least float whose correctly rounded square root
    // is at least tol, so that sqrtf(x) < tol exactly when x < stop2
    float stop2;
    // launch shape: subsets (waves) per fast-path workgroup (0 or 1: one),
    // and Path A's node order (0: row-major; T > 0: T x T tiles)
    int wpb, tile;
};

// Path A's counters are spread over slots of one 128-byte line each: one
// shared counter would take two same-address atomics per subset, which the L2
// serializes (about 4 ms of a 4096^2 frame's 262k subsets).
constexpr int kStatSlots = 512, kStatStride = 32;
// Sums the five counters of the slots into out[0..4] (after zeroing nothing:
// out is overwritten).
hipError_t launch_sum_stats(const int* slots, int* out, hipStream_t st);

// Pixels-per-lane class of a subset size (-1: larger than kMaxSubset).
int ppl_for(int dim);

// solver_*.hip: one translation unit per (interpolator, deformed pixel type).
// Launches the fast path over the work items, then the full solver over the
// items it handed on. a.bail must hold nitems ints; a.nbail is zeroed here.
// With a.defc set (and Keys 4x4 over an FP32 plane), the fast path's
// interior samples read the coefficient plane instead.
// phases: the fast path (with the hand-off counter's clear) and the full
// solver, or one of them (the full solver after the fast path of the same
// arguments on the same stream).
constexpr int kSolveFast = 1, kSolveFull = 2, kSolveBoth = 3;
hipError_t launch_solve(const SolveArgs& a, bool use6, bool def_f32, hipStream_t st, int phases = kSolveBoth);

// prep.hip
hipError_t launch_prep_ref_u8(const uint8_t* img, const uint8_t* mask, int w, int h, float* I, float* gx,
                              float* gy, hipStream_t st);
hipError_t launch_prep_ref_f32(const float* img, int w, int h, float* I, float* gx, float* gy,
                               hipStream_t st);
hipError_t launch_scan(const uint8_t* mask, int w, int h, int rect_x, int rect_y, int step, int gw, int gh,
                       int dim, uint8_t* skip, int* count, hipStream_t st, const int* dims = nullptr);
hipError_t launch_pool(const float* I, const float* gx, const float* gy, int w, int h, const uint8_t* skip,
                       int gw, int gh, int rect_x, int rect_y, int step, int dim, bool lm, PoolEntry* pool,
                       void* raw, hipStream_t st, const PoolOpts& opts = PoolOpts());
// Bytes of the scratch launch_pool needs for n grid nodes (the per-node sums).
size_t pool_raw_bytes(int n);
hipError_t launch_pool_list(const float* I, const float* gx, const float* gy, int w, int h, const int* cxy,
                            int n, int dim, bool lm, PoolEntry* pool, hipStream_t st);
hipError_t launch_downsample_u8(const uint8_t* src, int w, int h, float* dst, hipStream_t st);
// fp16 copy of a float plane (round to nearest), as half bits
hipError_t launch_to_f16(const float* src, size_t n, uint16_t* dst, hipStream_t st);
// bf16 copy of a float plane (round to nearest even)
hipError_t launch_to_bf16(const float* src, size_t n, uint16_t* dst, hipStream_t st);
// u8 -> FP32 copy of a plane (src 4-byte aligned)
hipError_t launch_u8_to_f32(const uint8_t* src, size_t n, float* dst, hipStream_t st);
// Row-cubic coefficient plane of a u8 frame: per pixel the Keys cubics of the
// cell to its right on its row and the next, four fp16 each (exact for u8
// taps); kCoefMaxPixels bounds the frame (32-bit byte offsets)
constexpr size_t kCoefMaxPixels = (size_t)1 << 27;
hipError_t launch_row_cubic(const uint8_t* img, int w, int h, uint4* out, hipStream_t st);
// solver_cf32.hip: the fast path's Keys 4x4 sampling forms at given points
// (0: FP32 plane, 1: coefficient plane, 2: u8 plane), for tests
hipError_t launch_interp_probe(const float* f32, const uint4* coef, const uint8_t* u8, int w, int h, int form,
                               int n, const float* xs, const float* ys, float* out, hipStream_t st);
hipError_t launch_downsample_f32(const float* src, int w, int h, float* dst, hipStream_t st);

// Global-motion patch search: per patch the best integer shift after the
// coarse-to-fine search, its ZNCC, the template std and the four ZNCC values
// around it (for the sub-pixel parabolas). Levels: 0 reads the u8 planes,
// 1..L the float pyramids.
struct PyrDev {
    const uint8_t* r0;
    const uint8_t* d0;
    const float* r[4];
    const float* d[4];
    int w[4], h[4];
    int levels;
    int coarse_radius;   // shifts searched at the coarsest level
    // The refinement after the rotation-invariant search: the deformed window
    // is sampled bilinearly through the linear map f (row-major 2 x 2: a
    // template offset o lands at shift + f o), and each patch's search starts
    // at start[2 k], start[2 k + 1] (coarsest-level pixels). warp = 0: f is
    // the identity and every search starts at zero shift.
    int warp;
    double f[4];
    const int* start;
};
// Coarse search radii of the global-motion patch search: the first pass, and
// the wider pass run when the first one's model does not lock.
constexpr int kCoarseRadius = 8, kCoarseRadiusWide = 16;
struct PatchOut {
    int sx, sy, ok, pad;
    double best, tstd, xm, xp, ym, yp;
    int has_x, has_y, pad2[2];
};
hipError_t launch_patch_search(const PyrDev& p, const int* centres, int npatch, PatchOut* out, void* scratch,
                               hipStream_t st);
// Bytes of the scratch launch_patch_search needs for npatch patches (the
// coarsest level's candidates of the blocks that share a patch; with a null
// scratch one block searches each patch alone).
size_t patch_scratch_bytes(int npatch);

// Rotation-invariant patch matching (the global-motion search for motions the
// translation search cannot follow): each reference patch and every position
// of the deformed image at one pyramid level is described by its disc turned
// to the orientation of its intensity centroid and sampled on a polar grid;
// a patch's match is the position whose description correlates best.
constexpr int kCanonR = 16;                          // disc radius (level pixels)
constexpr int kCanonRings = kCanonR - 2;             // sampled radii 2 .. 15
constexpr int kCanonAngles = 32;
constexpr int kCanonSamples = kCanonRings * kCanonAngles;   // 448
constexpr int kCanonMaxPatches = 25;
constexpr int kCanonMargin = kCanonR + 2;            // candidate positions keep the disc inside
struct CanonOut {
    int bx, by;          // best position (level pixels), -1 if none
    float best, second;  // its score, and the best one more than 2 px away
    float theta_ref, theta_def;   // the centroid orientations of the pair (rad)
    int pad[2];
};
// Descriptors of npatch reference patches (centres at level l), then the
// scores of every deformed position and the best per patch. desc holds
// npatch * kCanonSamples floats, scores (w_l - 2 m) * (h_l - 2 m) * npatch.
hipError_t launch_canon_match(const PyrDev& p, int l, const int* centres_l, int npatch, float* desc,
                              float* theta_ref, float* scores, CanonOut* out, hipStream_t st);
size_t canon_scores_floats(int wl, int hl, int npatch);

// guess.hip
struct GuessArgs {
    const TriDev* tris;
    const int* cell_tri;      // first triangle of lattice cell c: cell_tri[c] .. cell_tri[c + 1]
    const int* lat_x;         // lattice grid columns / rows
    const int* lat_y;
    int LW, LH;
    int quality;              // 2 = FULL (extrapolation on)
    int have_affine;
    double aff[6];
    double acx, acy;
    float gU, gV;
    int gw, gh, rect_x, rect_y, step;
};
hipError_t launch_guess(const GuessArgs& a, GridDev g, uint8_t* in_mesh, float* gu, float* gv, float* gux,
                        float* guy, float* gvx, float* gvy, float* tmp, hipStream_t st);

// strain.hip
// one median-test pass; with prev_fail non-null and *prev_fail == 0 the pass
// copies the flags (a pass after one that rejected nothing changes nothing)
hipError_t launch_outlier_pass(int gw, int gh, const float* u, const float* v, const uint8_t* ok_in,
                               uint8_t* ok_out, const int* prev_fail, int* nfail, hipStream_t st);
// Row-major pack of the surviving nodes into out (8 floats per point) with the
// capacity rule; counts[0] = points to write before the capacity cut (all
// strain-valid nodes), counts[1] += strain failures before the cut (counts[1]
// must be zero on entry). block_w holds pack_blocks(gw * gh) ints.
int pack_blocks(int n);
hipError_t launch_pack(int gw, int gh, int rect_x, int rect_y, int step, int cap_points, const uint8_t* ok,
                       const float* u, const float* v, const float* exx, const float* eyy, const float* exy,
                       const float* corr, int* block_w, int* counts, float* out, hipStream_t st);
hipError_t launch_vsg(int gw, int gh, int step, int window, int expected, const float* u, const float* v,
                      const uint8_t* valid, float* exx, float* eyy, float* exy, hipStream_t st);

// grid.hip (per-node bookkeeping)
hipError_t launch_grid_init(GridDev g, const uint8_t* skip, int n, hipStream_t st);
// Incremental-session anchor patches: n entries (slot, x0, y0) cut side x side
// blocks of a w-wide u8 frame into a slot store; and the matching canvas, the
// store slots of a list placed in rows of `cols` cells on a zeroed w x
// (rows side) canvas.
hipError_t launch_cut_patches(const uint8_t* frame, int w, const int32_t* entries, int n, int side, uint8_t* store,
                              hipStream_t st);
hipError_t launch_place_patches(const uint8_t* store, int side, const int32_t* slots, int n, int cols, int w,
                                int rows, uint8_t* canvas, hipStream_t st);
// Write solve results into the grid: with accepted_only, only results with
// status 0 and corr <= kCorrAccept mark their node solved (Path B commits);
// otherwise every listed result does (published anchors, with tag).
hipError_t launch_apply(GridDev g, const int* nodes, const SolveOut* res, int n, int tag, bool accepted_only,
                        hipStream_t st);
// valid = solved && corr >= 0
hipError_t launch_valid(GridDev g, uint8_t* valid, int n, hipStream_t st);
// wave_sum and wave_sum_i of nw waves of 64 inputs, every lane's result
hipError_t launch_wave_sum_probe(const float* in, int nw, float* out, int* iout, hipStream_t st);
// n ints set to zero by a fill kernel (no memset node in a captured graph)
hipError_t launch_zero(int* p, int n, hipStream_t st);
// count of nodes with solved == 0 into *count (zeroed first by the launch)
hipError_t launch_count_open(GridDev g, int n, int* count, hipStream_t st);
}  // namespace sdic
