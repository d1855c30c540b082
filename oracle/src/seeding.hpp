// CPU ORACLE — test infrastructure only (see oracle.hpp).
// Seeding front end. The reference fits an AKAZE homography (OpenCV, absent
// here) and triangulates the kept anchors with cv::Subdiv2D. Both engines of
// this project replace those with:
//   * a coarse-to-fine ZNCC patch search on 2x2 box pyramids, fitted to an
//     affine model (the "global model": lock, translation at the ROI centre,
//     gradients);
//   * a triangulation of the anchor lattice itself (two triangles per lattice
//     cell with four kept corners, one with three);
//   * a robust affine refit to the kept anchors in place of the RANSAC
//     homography refit.
// Everything else of solve_anchor_seeds / build_mesh_guess_field is restated.
#pragma once

#include <cstdint>
#include <vector>

namespace orc {

struct Pyramid {
    std::vector<std::vector<float>> lv;
    std::vector<int> w, h;
};
void build_pyramid(Pyramid& p, const float* img, int w, int h, int levels);
int pyramid_levels(int rect_w, int rect_h);

struct GlobalModel {
    bool locked = false;
    bool full = false;
    double a[6] = {0, 0, 0, 0, 0, 0};  // u = a0 + a1 (x-cx) + a2 (y-cy); v = a3 + a4 (.) + a5 (.)
    double cx = 0, cy = 0;
    double t_u = 0, t_v = 0;
    int patches = 0, inliers = 0;
    void at(double x, double y, float g[6]) const;
};
GlobalModel global_motion(const Pyramid& ref, const Pyramid& def, int rect_x, int rect_y,
                          int rect_w, int rect_h);

// Robust affine fit (iterated 1 px rejection) to point pairs; false if fewer
// than min_pts survive. Same parametrisation as GlobalModel, about (cx, cy).
bool robust_affine(const std::vector<double>& xr, const std::vector<double>& yr,
                   const std::vector<double>& ud, const std::vector<double>& vd, double cx,
                   double cy, int min_pts, double tol, double a[6]);

// Lattice triangulation: triangles over kept lattice nodes.
struct Tri {
    float px[3], py[3];            // reference positions
    double ux, uy, vx, vy, u, v;   // affine: d = (ux x + uy y + u, vx x + vy y + v) - (x, y)
    float bx0, by0, bx1, by1;      // bounding box +-15 px (full_field_mesh.cpp:131)
    int cell;                      // lattice cell index lj * (LW-1) + li
};
void lattice_triangles(int LW, int LH, const std::vector<unsigned char>& kept,
                       const std::vector<float>& rx, const std::vector<float>& ry,
                       const std::vector<float>& u, const std::vector<float>& v,
                       std::vector<Tri>& tris);
// pointPolygonTest semantics over a triangle: > 0 inside, 0 on an edge, < 0
// outside; with measure, the signed distance to the nearest edge.
double tri_test(const Tri& t, float x, float y, bool measure);

}  // namespace orc
