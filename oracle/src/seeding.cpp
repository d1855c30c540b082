// CPU ORACLE — test infrastructure only (see oracle.hpp and seeding.hpp).
#include "seeding.hpp"

#include <algorithm>
#include <cmath>

#include "oracle.hpp"

namespace orc {

void build_pyramid(Pyramid& p, const float* img, int w, int h, int levels) {
    p.lv.assign(levels + 1, {});
    p.w.assign(levels + 1, 0);
    p.h.assign(levels + 1, 0);
    p.lv[0].assign(img, img + (size_t)w * h);
    p.w[0] = w;
    p.h[0] = h;
    for (int l = 1; l <= levels; ++l) {
        const int pw = p.w[l - 1], nw = pw / 2, nh = p.h[l - 1] / 2;
        p.w[l] = nw;
        p.h[l] = nh;
        p.lv[l].resize((size_t)nw * nh);
        const float* s = p.lv[l - 1].data();
        for (int y = 0; y < nh; ++y)
            for (int x = 0; x < nw; ++x) {
                const float* a = s + (size_t)(2 * y) * pw + 2 * x;
                p.lv[l][(size_t)y * nw + x] = 0.25f * ((a[0] + a[1]) + (a[pw] + a[pw + 1]));
            }
    }
}

int pyramid_levels(int rect_w, int rect_h) {
    const int m = std::min(rect_w, rect_h);
    int L = 0;
    while (L < 3 && (m >> (L + 1)) >= 64) ++L;
    return L;
}

void GlobalModel::at(double x, double y, float g[6]) const {
    const double dx = x - cx, dy = y - cy;
    g[0] = (float)(a[0] + a[1] * dx + a[2] * dy);
    g[1] = (float)(a[3] + a[4] * dx + a[5] * dy);
    g[2] = (float)a[1];
    g[3] = (float)a[2];
    g[4] = (float)a[4];
    g[5] = (float)a[5];
}

namespace {

constexpr int kPatchHalf = 16;     // 32 x 32 template at every level
constexpr int kCoarseRadius = 8;   // shifts at the coarsest level
constexpr int kCoarseRadiusWide = 16;   // the second pass when the first does not lock
constexpr int kFineRadius = 2;     // around twice the coarser estimate
// patch centres per side: 5, or 8 on ROIs of 2048 px and more
int patch_grid(int rect_w, int rect_h) { return std::min(rect_w, rect_h) >= 2048 ? 8 : 5; }
constexpr double kMinTemplateStd = 2.0;
constexpr double kMinScore = 0.5;
// consensus fit over random triples when the least squares does not hold half
// the patches
constexpr int kConsensusTrials = 600;
constexpr double kConsensusTol = 3.0;
constexpr int kConsensusMinInliers = 6;
constexpr double kConsensusMinArea = 1.0e4;

// ZNCC of the template at (cx, cy) against the window shifted by (sx, sy) at
// one pyramid level; false if either window leaves the image.
bool zncc(const Pyramid& R, const Pyramid& D, int l, int cx, int cy, int sx, int sy, double& out,
          double* tstd = nullptr) {
    const int w = R.w[l], h = R.h[l];
    const int x0 = cx - kPatchHalf, y0 = cy - kPatchHalf, x1 = cx + kPatchHalf, y1 = cy + kPatchHalf;
    if (x0 < 0 || y0 < 0 || x1 > w || y1 > h) return false;
    if (x0 + sx < 0 || y0 + sy < 0 || x1 + sx > w || y1 + sy > h) return false;
    const float* r = R.lv[l].data();
    const float* d = D.lv[l].data();
    double sr = 0, sd = 0, srr = 0, sdd = 0, srd = 0;
    for (int y = y0; y < y1; ++y)
        for (int x = x0; x < x1; ++x) {
            const double a = r[(size_t)y * w + x], b = d[(size_t)(y + sy) * w + x + sx];
            sr += a;
            sd += b;
            srr += a * a;
            sdd += b * b;
            srd += a * b;
        }
    const double n = (double)(4 * kPatchHalf * kPatchHalf);
    const double vr = srr - sr * sr / n, vd = sdd - sd * sd / n;
    if (tstd) *tstd = std::sqrt(std::max(0.0, vr / n));
    if (!(vr > 1e-9) || !(vd > 1e-9)) {
        out = -2.0;
        return true;
    }
    out = (srd - sr * sd / n) / (std::sqrt(vr) * std::sqrt(vd));
    return true;
}

// Best shift in [c - r, c + r]^2: highest score, ties to the smaller |shift|.
bool search(const Pyramid& R, const Pyramid& D, int l, int cx, int cy, int csx, int csy, int rad,
            int& bx, int& by, double& best) {
    bool any = false;
    best = -3.0;
    long br2 = 0;
    for (int sy = csy - rad; sy <= csy + rad; ++sy)
        for (int sx = csx - rad; sx <= csx + rad; ++sx) {
            double s;
            if (!zncc(R, D, l, cx, cy, sx, sy, s)) continue;
            const long r2 = (long)sx * sx + (long)sy * sy;
            if (!any || s > best || (s == best && r2 < br2)) {
                any = true;
                best = s;
                bx = sx;
                by = sy;
                br2 = r2;
            }
        }
    return any;
}

bool solve3(const double A[9], const double b[3], double x[3]) {
    auto M = [&](int r, int c) { return A[r * 3 + c]; };
    auto cof = [&](int i, int j) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        return M(i1, j1) * M(i2, j2) - M(i1, j2) * M(i2, j1);
    };
    const double det = cof(0, 0) * M(0, 0) + cof(1, 0) * M(1, 0) + cof(2, 0) * M(2, 0);
    if (!(std::fabs(det) > 1e-12) || !finite_bits(det)) return false;
    for (int i = 0; i < 3; ++i) {
        double s = 0;
        for (int j = 0; j < 3; ++j) s += cof(j, i) * b[j];
        x[i] = s / det;
    }
    return true;
}

}  // namespace

bool robust_affine(const std::vector<double>& xr, const std::vector<double>& yr,
                   const std::vector<double>& ud, const std::vector<double>& vd, double cx,
                   double cy, int min_pts, double tol, double a[6]) {
    const size_t n = xr.size();
    std::vector<unsigned char> in(n, 1);
    bool have = false;
    for (int it = 0; it < 5; ++it) {
        double A[9] = {}, bu[3] = {}, bv[3] = {};
        int m = 0;
        for (size_t i = 0; i < n; ++i) {
            if (!in[i]) continue;
            const double r[3] = {1.0, xr[i] - cx, yr[i] - cy};
            for (int p = 0; p < 3; ++p) {
                for (int q = 0; q < 3; ++q) A[p * 3 + q] += r[p] * r[q];
                bu[p] += r[p] * ud[i];
                bv[p] += r[p] * vd[i];
            }
            ++m;
        }
        if (m < min_pts) return false;
        double su[3], sv[3];
        if (!solve3(A, bu, su) || !solve3(A, bv, sv)) return false;
        a[0] = su[0];
        a[1] = su[1];
        a[2] = su[2];
        a[3] = sv[0];
        a[4] = sv[1];
        a[5] = sv[2];
        have = true;
        bool changed = false;
        int kept = 0;
        for (size_t i = 0; i < n; ++i) {
            const double dx = xr[i] - cx, dy = yr[i] - cy;
            const double eu = a[0] + a[1] * dx + a[2] * dy - ud[i];
            const double ev = a[3] + a[4] * dx + a[5] * dy - vd[i];
            const unsigned char keep = std::sqrt(eu * eu + ev * ev) <= tol;
            if (keep != in[i]) changed = true;
            in[i] = keep;
            kept += keep;
        }
        if (kept < min_pts) return false;
        if (!changed) break;
    }
    return have;
}

// Deterministic consensus affine: triples drawn by a fixed LCG; the affine
// through the triple holding the most patches within kConsensusTol (ties: the
// smaller summed residual); refined by robust_affine on that set.
bool consensus_affine(const std::vector<double>& x, const std::vector<double>& y,
                      const std::vector<double>& u, const std::vector<double>& v, double cx, double cy,
                      double a[6]) {
    const int n = (int)x.size();
    const int need = std::max(kConsensusMinInliers, n / 4);
    if (n < need || n < 3) return false;
    uint64_t lcg = 0x9E3779B97F4A7C15ull;
    auto draw = [&](int m) {
        lcg = lcg * 6364136223846793005ull + 1442695040888963407ull;
        return (int)((lcg >> 33) % (uint64_t)m);
    };
    int best_in = 0;
    double best_res = 0.0, best[6] = {0, 0, 0, 0, 0, 0};
    for (int t = 0; t < kConsensusTrials; ++t) {
        const int i = draw(n), j = draw(n), k = draw(n);
        if (i == j || j == k || i == k) continue;
        const double A[9] = {1.0, x[i] - cx, y[i] - cy, 1.0, x[j] - cx, y[j] - cy, 1.0, x[k] - cx, y[k] - cy};
        const double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                           A[2] * (A[3] * A[7] - A[4] * A[6]);
        if (!(std::fabs(det) > kConsensusMinArea)) continue;
        const double bu[3] = {u[i], u[j], u[k]}, bv[3] = {v[i], v[j], v[k]};
        double su[3], sv[3];
        if (!solve3(A, bu, su) || !solve3(A, bv, sv)) continue;
        int in = 0;
        double res = 0.0;
        for (int q = 0; q < n; ++q) {
            const double dx = x[q] - cx, dy = y[q] - cy;
            const double e = std::hypot(su[0] + su[1] * dx + su[2] * dy - u[q], sv[0] + sv[1] * dx + sv[2] * dy - v[q]);
            if (e <= kConsensusTol) {
                ++in;
                res += e;
            }
        }
        if (in > best_in || (in == best_in && in > 0 && res < best_res)) {
            best_in = in;
            best_res = res;
            for (int q = 0; q < 3; ++q) {
                best[q] = su[q];
                best[3 + q] = sv[q];
            }
        }
    }
    if (best_in < need) return false;
    std::vector<double> xi, yi, ui, vi;
    for (int q = 0; q < n; ++q) {
        const double dx = x[q] - cx, dy = y[q] - cy;
        const double e = std::hypot(best[0] + best[1] * dx + best[2] * dy - u[q], best[3] + best[4] * dx + best[5] * dy - v[q]);
        if (e <= kConsensusTol) {
            xi.push_back(x[q]);
            yi.push_back(y[q]);
            ui.push_back(u[q]);
            vi.push_back(v[q]);
        }
    }
    return robust_affine(xi, yi, ui, vi, cx, cy, need, kConsensusTol, a);
}

GlobalModel global_motion_pass(const Pyramid& ref, const Pyramid& def, int rect_x, int rect_y, int rect_w,
                               int rect_h, int coarse_radius);

// The turn of a model's linear part, atan2(F10 - F01, F00 + F11) in degrees.
// A translation pass's untouched templates decorrelate long before a quarter
// turn, so a model it fits with more than 30 degrees is a false lock.
static bool plausible(const GlobalModel& gm) {
    const double f00 = 1.0 + gm.a[1], f01 = gm.a[2], f10 = gm.a[4], f11 = 1.0 + gm.a[5];
    return std::fabs(std::atan2(f10 - f01, f00 + f11) * (180.0 / 3.14159265358979323846)) <= 30.0;
}

GlobalModel global_motion(const Pyramid& ref, const Pyramid& def, int rect_x, int rect_y,
                          int rect_w, int rect_h) {
    GlobalModel gm = global_motion_pass(ref, def, rect_x, rect_y, rect_w, rect_h, kCoarseRadius);
    if (gm.locked && !plausible(gm)) gm = GlobalModel{};
    if (!gm.locked) gm = global_motion_pass(ref, def, rect_x, rect_y, rect_w, rect_h, kCoarseRadiusWide);
    if (gm.locked && !plausible(gm)) gm = GlobalModel{};
    return gm;
}

GlobalModel global_motion_pass(const Pyramid& ref, const Pyramid& def, int rect_x, int rect_y, int rect_w,
                               int rect_h, int coarse_radius) {
    GlobalModel gm;
    gm.cx = rect_x + 0.5 * rect_w;
    gm.cy = rect_y + 0.5 * rect_h;
    const int L = (int)ref.lv.size() - 1;
    std::vector<double> xs, ys, us, vs;
    const int grid = patch_grid(rect_w, rect_h);
    for (int j = 0; j < grid; ++j)
        for (int i = 0; i < grid; ++i) {
            const int px = rect_x + (int)std::lround((i + 0.5) * rect_w / grid);
            const int py = rect_y + (int)std::lround((j + 0.5) * rect_h / grid);
            int sx = 0, sy = 0;
            double best = -3.0;
            bool ok = true;
            for (int l = L; l >= 0 && ok; --l) {
                const int rad = (l == L) ? coarse_radius : kFineRadius;
                int bx = 0, by = 0;
                ok = search(ref, def, l, px >> l, py >> l, sx, sy, rad, bx, by, best);
                sx = bx;
                sy = by;
                if (l > 0) {
                    sx *= 2;
                    sy *= 2;
                }
            }
            if (!ok) continue;
            double tstd = 0, s0;
            zncc(ref, def, 0, px, py, sx, sy, s0, &tstd);
            if (!(tstd > kMinTemplateStd) || !(best >= kMinScore)) continue;
            // sub-pixel: 1-D parabolas through the neighbouring integer shifts
            double fx = 0, fy = 0, a, b;
            if (zncc(ref, def, 0, px, py, sx - 1, sy, a) && zncc(ref, def, 0, px, py, sx + 1, sy, b)) {
                const double den = a - 2 * best + b;
                if (den < 0) fx = std::max(-0.5, std::min(0.5, 0.5 * (a - b) / den));
            }
            if (zncc(ref, def, 0, px, py, sx, sy - 1, a) && zncc(ref, def, 0, px, py, sx, sy + 1, b)) {
                const double den = a - 2 * best + b;
                if (den < 0) fy = std::max(-0.5, std::min(0.5, 0.5 * (a - b) / den));
            }
            xs.push_back(px);
            ys.push_back(py);
            us.push_back(sx + fx);
            vs.push_back(sy + fy);
        }
    gm.patches = (int)xs.size();
    if (gm.patches < 3) return gm;
    double a[6];
    const int need = std::max(3, (gm.patches + 1) / 2);
    if (!robust_affine(xs, ys, us, vs, gm.cx, gm.cy, need, 2.0, a) &&
        !consensus_affine(xs, ys, us, vs, gm.cx, gm.cy, a))
        return gm;
    for (int k = 0; k < 6; ++k) gm.a[k] = a[k];
    gm.locked = true;
    gm.t_u = a[0];
    gm.t_v = a[3];
    int inl = 0;
    for (size_t i = 0; i < xs.size(); ++i) {
        const double dx = xs[i] - gm.cx, dy = ys[i] - gm.cy;
        const double eu = a[0] + a[1] * dx + a[2] * dy - us[i], ev = a[3] + a[4] * dx + a[5] * dy - vs[i];
        inl += std::sqrt(eu * eu + ev * ev) <= 2.0;
    }
    gm.inliers = inl;
    const double cxs[2] = {(double)rect_x, (double)(rect_x + rect_w)};
    const double cys[2] = {(double)rect_y, (double)(rect_y + rect_h)};
    for (double x : cxs)
        for (double y : cys) {
            float g[6];
            gm.at(x, y, g);
            if (std::hypot(g[0] - gm.t_u, g[1] - gm.t_v) > kSimilarityMinCornerPx) gm.full = true;
        }
    return gm;
}

void lattice_triangles(int LW, int LH, const std::vector<unsigned char>& kept,
                       const std::vector<float>& rx, const std::vector<float>& ry,
                       const std::vector<float>& u, const std::vector<float>& v,
                       std::vector<Tri>& tris) {
    tris.clear();
    for (int lj = 0; lj + 1 < LH; ++lj)
        for (int li = 0; li + 1 < LW; ++li) {
            const int c[4] = {lj * LW + li, lj * LW + li + 1, (lj + 1) * LW + li + 1,
                              (lj + 1) * LW + li};
            int nk = 0;
            for (int k = 0; k < 4; ++k) nk += kept[c[k]] ? 1 : 0;
            int sets[2][3];
            int ns = 0;
            if (nk == 4) {
                sets[0][0] = c[0], sets[0][1] = c[1], sets[0][2] = c[2];
                sets[1][0] = c[0], sets[1][1] = c[2], sets[1][2] = c[3];
                ns = 2;
            } else if (nk == 3) {
                int m = 0;
                for (int k = 0; k < 4; ++k)
                    if (kept[c[k]]) sets[0][m++] = c[k];
                ns = 1;
            }
            for (int s = 0; s < ns; ++s) {
                Tri t;
                t.cell = lj * (LW - 1) + li;
                double A[9], bx[3], by[3];
                for (int k = 0; k < 3; ++k) {
                    const int q = sets[s][k];
                    t.px[k] = rx[q];
                    t.py[k] = ry[q];
                    const float dxp = rx[q] + u[q], dyp = ry[q] + v[q];
                    A[k * 3] = t.px[k];
                    A[k * 3 + 1] = t.py[k];
                    A[k * 3 + 2] = 1.0;
                    bx[k] = dxp;
                    by[k] = dyp;
                }
                double mx[3], my[3];
                if (!solve3(A, bx, mx) || !solve3(A, by, my)) continue;
                const double detF = mx[0] * my[1] - mx[1] * my[0];
                if (!finite_bits(detF) || std::fabs(detF) <= 1e-3) continue;
                t.ux = mx[0] - 1.0;
                t.uy = mx[1];
                t.u = mx[2];
                t.vx = my[0];
                t.vy = my[1] - 1.0;
                t.v = my[2];
                const float x0 = std::min({t.px[0], t.px[1], t.px[2]}),
                            x1 = std::max({t.px[0], t.px[1], t.px[2]});
                const float y0 = std::min({t.py[0], t.py[1], t.py[2]}),
                            y1 = std::max({t.py[0], t.py[1], t.py[2]});
                t.bx0 = x0 - 15.0f;
                t.by0 = y0 - 15.0f;
                t.bx1 = x1 + 15.0f;
                t.by1 = y1 + 15.0f;
                tris.push_back(t);
            }
        }
}

double tri_test(const Tri& t, float x, float y, bool measure) {
    int pos = 0, neg = 0, zero = 0;
    double dmin = 1e300;
    for (int k = 0; k < 3; ++k) {
        const double ax = t.px[k], ay = t.py[k];
        const double bx = t.px[(k + 1) % 3], by = t.py[(k + 1) % 3];
        const double cr = (bx - ax) * ((double)y - ay) - (by - ay) * ((double)x - ax);
        if (cr > 0) pos++;
        else if (cr < 0) neg++;
        else zero++;
        if (measure) {
            const double ex = bx - ax, ey = by - ay;
            const double l2 = ex * ex + ey * ey;
            double tt = l2 > 0 ? ((x - ax) * ex + (y - ay) * ey) / l2 : 0.0;
            tt = std::max(0.0, std::min(1.0, tt));
            const double qx = ax + tt * ex - x, qy = ay + tt * ey - y;
            dmin = std::min(dmin, std::sqrt(qx * qx + qy * qy));
        }
    }
    const bool inside = (pos == 0 || neg == 0);
    const bool on_edge = inside && zero > 0;
    if (!measure) return on_edge ? 0.0 : (inside ? 1.0 : -1.0);
    if (on_edge) return 0.0;
    return inside ? dmin : -dmin;
}

}  // namespace orc
