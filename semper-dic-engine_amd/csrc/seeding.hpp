// Host side of the seeding stage (integer / small-array work between kernels):
// the anchor-lattice plan and median test (full_field_anchors.cpp:64-168),
// convex-hull coverage, the global affine model fitted to the patch search,
// the robust refit to the kept anchors, and the lattice triangulation that
// stands in for the reference's Delaunay mesh.
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"
#include "solver.hpp"

namespace sdic {

int lattice_count(int n, int stride);
// plan_anchor_lattice (full_field_anchors.cpp:73-120) for about `target`
// anchors (the reference's kAnchorTarget, 256, unless the learned seeder asks
// for a denser lattice)
void plan_lattice(int gw, int gh, int& sx, int& sy, int target = kAnchorTarget);

// median_inplace (src/util/median.hpp:21-30)
float median_of(float* v, size_t n);

struct AnchorRec {  // AnchorResult (full_field_internal.hpp:422-433)
    int gx = 0, gy = 0, node = -1;
    float rx = 0, ry = 0, p[6] = {0, 0, 0, 0, 0, 0}, corr = 0;
    int iters = 0;
    bool simplex = false, vertex = false, output = false, kept = false;
};
// mark_kept_anchors (full_field_anchors.cpp:122-168)
void mark_kept(std::vector<AnchorRec>& an, int LW, int LH, const std::vector<float>& ru,
               const std::vector<float>& rv);

double hull_area(std::vector<std::pair<double, double>> p);

// Affine model u = a0 + a1 (x - cx) + a2 (y - cy), v = a3 + a4 (.) + a5 (.):
// least squares with iterated rejection at `tol` px; false below min_pts.
bool robust_affine(const std::vector<double>& x, const std::vector<double>& y, const std::vector<double>& u,
                   const std::vector<double>& v, double cx, double cy, int min_pts, double tol, double a[6]);

// The same model from a consensus of random triples (tolerance 3 px, at least
// max(6, n / 4) patches), refined by robust_affine on the consensus set.
bool consensus_affine(const std::vector<double>& x, const std::vector<double>& y, const std::vector<double>& u,
                      const std::vector<double>& v, double cx, double cy, double a[6],
                      double tol = 3.0);

struct GlobalModel {
    bool locked = false, full = false;
    double a[6] = {0, 0, 0, 0, 0, 0};
    double cx = 0, cy = 0, t_u = 0, t_v = 0;
    int patches = 0;
    void at(double x, double y, float g[6]) const;
};
int pyramid_levels(int rect_w, int rect_h);
void patch_centres(int rect_x, int rect_y, int rect_w, int rect_h, std::vector<int>& cxy);
GlobalModel fit_global_model(const std::vector<int>& cxy, const std::vector<PatchOut>& po, int rect_x,
                             int rect_y, int rect_w, int rect_h);
// The turn of a model's linear part F = I + grad u, in degrees: the angle of
// its rotation factor, atan2(F10 - F01, F00 + F11).
double model_turn_deg(const GlobalModel& gm);
// The largest turn a translation pass's model may carry; beyond it the model
// is a false lock and the rotation-invariant pass decides.
constexpr double kTranslationMaxTurnDeg = 30.0;
// The rough affine of the rotation-invariant pass: the matches of the patch
// centres (level-l pixels, CanonOut) with a clear best (score >= 0.5, 0.05
// above any position more than 2 px away), in full-resolution pixels (a
// level-l pixel i is centred on i 2^l + (2^l - 1) / 2), through the consensus
// fit with a tolerance of 2^(l+1) px. False when it does not hold.
bool canon_affine(const std::vector<int>& centres_l, const std::vector<CanonOut>& co, int l, double cx, double cy,
                  double a[6]);

// Two triangles per lattice cell with four kept corners, one with three, in
// cell order; cell_tri[c] .. cell_tri[c + 1] are cell c's.
void lattice_triangles(int LW, int LH, const std::vector<unsigned char>& kept, const std::vector<float>& rx,
                       const std::vector<float>& ry, const std::vector<float>& u, const std::vector<float>& v,
                       std::vector<TriDev>& tris, std::vector<int>& cell_tri);

}  // namespace sdic
