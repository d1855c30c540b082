// Host side of the seeding stage; see seeding.hpp.
#include "seeding.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace sdic {

namespace {

bool finite_d(double d) {
    uint64_t u;
    std::memcpy(&u, &d, sizeof u);
    return (u & 0x7ff0000000000000ull) != 0x7ff0000000000000ull;
}

// 3x3 solve by cofactors (det from the first column); false when singular.
bool solve3(const double A[9], const double b[3], double x[3]) {
    auto M = [&](int r, int c) { return A[r * 3 + c]; };
    auto cof = [&](int i, int j) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        return M(i1, j1) * M(i2, j2) - M(i1, j2) * M(i2, j1);
    };
    const double det = cof(0, 0) * M(0, 0) + cof(1, 0) * M(1, 0) + cof(2, 0) * M(2, 0);
    if (!(std::fabs(det) > 1e-12) || !finite_d(det)) return false;
    for (int i = 0; i < 3; ++i) {
        double s = 0;
        for (int j = 0; j < 3; ++j) s += cof(j, i) * b[j];
        x[i] = s / det;
    }
    return true;
}

// patch centres per side: 5, or 8 on ROIs of 2048 px and more (where local
// motion has room to stray from one affine)
int patch_grid(int rect_w, int rect_h) { return std::min(rect_w, rect_h) >= 2048 ? 8 : 5; }
constexpr double kMinTemplateStd = 2.0;
constexpr double kMinScore = 0.5;
// The consensus fit when the plain least squares does not hold half the
// patches (local motion beyond an affine): deterministic random triples, the
// affine through the one with the most patches within its tolerance (3 px for
// the translation passes' shifts; seeding.hpp).
constexpr int kConsensusTrials = 600;
constexpr int kConsensusMinInliers = 6;
constexpr double kConsensusMinArea = 1.0e4;   // |det| of the triple's design matrix, px^2

}  // namespace

int lattice_count(int n, int stride) {
    if (n <= 0 || stride <= 0) return 0;
    int c = (n + stride - 1) / stride;
    if ((c - 1) * stride != n - 1) c++;
    return c;
}

void plan_lattice(int gw, int gh, int& sx, int& sy, int target) {
    int stride = (int)std::lround(std::sqrt((double)(gw * gh) / (double)target));
    stride = std::max(1, stride);
    const int budget = lattice_count(gw, stride) * lattice_count(gh, stride);
    sx = sy = stride;
    while (sx > 1 && lattice_count(gw, sx) < 3) --sx;
    while (sy > 1 && lattice_count(gh, sy) < 3) --sy;
    while (lattice_count(gw, sx) * lattice_count(gh, sy) > budget) {
        const bool grow_x = lattice_count(gw, sx + 1) >= 3;
        const bool grow_y = lattice_count(gh, sy + 1) >= 3;
        if (!grow_x && !grow_y) break;
        if (grow_x && (!grow_y || lattice_count(gw, sx) >= lattice_count(gh, sy))) ++sx;
        else ++sy;
    }
}

float median_of(float* v, size_t n) {
    n = (size_t)(std::partition(v, v + n, [](float x) { return x == x; }) - v);
    if (n == 0) return 0.0f;
    const size_t mid = n / 2;
    std::nth_element(v, v + mid, v + n);
    const float hi = v[mid];
    if (n % 2 == 1) return hi;
    return 0.5f * (*std::max_element(v, v + mid) + hi);
}

void mark_kept(std::vector<AnchorRec>& an, int LW, int LH, const std::vector<float>& ru,
               const std::vector<float>& rv) {
    for (int lj = 0; lj < LH; ++lj)
        for (int li = 0; li < LW; ++li) {
            const size_t i = (size_t)lj * LW + li;
            AnchorRec& a = an[i];
            a.kept = false;
            if (!a.vertex) continue;
            float nu[8], nv[8];
            int nn = 0;
            for (int dj = -1; dj <= 1; ++dj)
                for (int di = -1; di <= 1; ++di) {
                    if (!di && !dj) continue;
                    const int nj = lj + dj, ni = li + di;
                    if (ni < 0 || ni >= LW || nj < 0 || nj >= LH) continue;
                    const size_t q = (size_t)nj * LW + ni;
                    if (!an[q].vertex) continue;
                    nu[nn] = ru[q];
                    nv[nn] = rv[q];
                    ++nn;
                }
            if (nn < 3) continue;
            const float mu = median_of(nu, nn), mv = median_of(nv, nn);
            if (std::fabs(ru[i] - mu) > kAnchorMedianTol || std::fabs(rv[i] - mv) > kAnchorMedianTol) continue;
            a.kept = true;
        }
}

double hull_area(std::vector<std::pair<double, double>> p) {
    if (p.size() < 3) return 0.0;
    std::sort(p.begin(), p.end());
    p.erase(std::unique(p.begin(), p.end()), p.end());
    if (p.size() < 3) return 0.0;
    auto cross = [](const std::pair<double, double>& o, const std::pair<double, double>& a,
                    const std::pair<double, double>& b) {
        return (a.first - o.first) * (b.second - o.second) - (a.second - o.second) * (b.first - o.first);
    };
    std::vector<std::pair<double, double>> h(2 * p.size());
    size_t k = 0;
    for (size_t i = 0; i < p.size(); ++i) {
        while (k >= 2 && cross(h[k - 2], h[k - 1], p[i]) <= 0) k--;
        h[k++] = p[i];
    }
    for (size_t i = p.size() - 1, t = k + 1; i > 0; --i) {
        while (k >= t && cross(h[k - 2], h[k - 1], p[i - 1]) <= 0) k--;
        h[k++] = p[i - 1];
    }
    h.resize(k - 1);
    double a = 0;
    for (size_t i = 0; i < h.size(); ++i) {
        const auto& q = h[i];
        const auto& r = h[(i + 1) % h.size()];
        a += q.first * r.second - r.first * q.second;
    }
    return std::fabs(a) * 0.5;
}

bool robust_affine(const std::vector<double>& x, const std::vector<double>& y, const std::vector<double>& u,
                   const std::vector<double>& v, double cx, double cy, int min_pts, double tol, double a[6]) {
    const size_t n = x.size();
    std::vector<unsigned char> in(n, 1);
    bool have = false;
    for (int it = 0; it < 5; ++it) {
        double A[9] = {}, bu[3] = {}, bv[3] = {};
        int m = 0;
        for (size_t i = 0; i < n; ++i) {
            if (!in[i]) continue;
            const double r[3] = {1.0, x[i] - cx, y[i] - cy};
            for (int p = 0; p < 3; ++p) {
                for (int q = 0; q < 3; ++q) A[p * 3 + q] += r[p] * r[q];
                bu[p] += r[p] * u[i];
                bv[p] += r[p] * v[i];
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
            const double dx = x[i] - cx, dy = y[i] - cy;
            const double eu = a[0] + a[1] * dx + a[2] * dy - u[i];
            const double ev = a[3] + a[4] * dx + a[5] * dy - v[i];
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

bool consensus_affine(const std::vector<double>& x, const std::vector<double>& y, const std::vector<double>& u,
                      const std::vector<double>& v, double cx, double cy, double a[6],
                      double tol) {
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
        // a triple spanning too little area fixes no gradient
        if (!(std::fabs(det) > kConsensusMinArea)) continue;
        const double bu[3] = {u[i], u[j], u[k]}, bv[3] = {v[i], v[j], v[k]};
        double su[3], sv[3];
        if (!solve3(A, bu, su) || !solve3(A, bv, sv)) continue;
        int in = 0;
        double res = 0.0;
        for (int q = 0; q < n; ++q) {
            const double dx = x[q] - cx, dy = y[q] - cy;
            const double eu = su[0] + su[1] * dx + su[2] * dy - u[q], ev = sv[0] + sv[1] * dx + sv[2] * dy - v[q];
            // a patch clearly beyond the tolerance is out without the hypot
            // (which is within an ulp of the root, far inside the margin)
            if (eu * eu + ev * ev > tol * tol * (1.0 + 1e-9)) continue;
            const double e = std::hypot(eu, ev);
            if (e <= tol) {
                ++in;
                res += e;
            }
        }
        if (in > best_in || (in == best_in && in > 0 && res < best_res)) {
            best_in = in;
            best_res = res;
            best[0] = su[0];
            best[1] = su[1];
            best[2] = su[2];
            best[3] = sv[0];
            best[4] = sv[1];
            best[5] = sv[2];
        }
    }
    if (best_in < need) return false;
    // least squares on the consensus set, then the iterated rejection
    std::vector<double> xi, yi, ui, vi;
    for (int q = 0; q < n; ++q) {
        const double dx = x[q] - cx, dy = y[q] - cy;
        const double e = std::hypot(best[0] + best[1] * dx + best[2] * dy - u[q],
                                    best[3] + best[4] * dx + best[5] * dy - v[q]);
        if (e <= tol) {
            xi.push_back(x[q]);
            yi.push_back(y[q]);
            ui.push_back(u[q]);
            vi.push_back(v[q]);
        }
    }
    return robust_affine(xi, yi, ui, vi, cx, cy, need, tol, a);
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

double model_turn_deg(const GlobalModel& gm) {
    const double f00 = 1.0 + gm.a[1], f01 = gm.a[2], f10 = gm.a[4], f11 = 1.0 + gm.a[5];
    return std::atan2(f10 - f01, f00 + f11) * (180.0 / 3.14159265358979323846);
}

bool canon_affine(const std::vector<int>& centres_l, const std::vector<CanonOut>& co, int l, double cx, double cy,
                  double a[6]) {
    const double s = (double)(1 << l), half = 0.5 * (s - 1.0);
    std::vector<double> xs, ys, us, vs;
    for (size_t k = 0; k < co.size(); ++k) {
        const CanonOut& o = co[k];
        if (o.bx < 0 || !(o.best >= 0.5f) || !(o.best - o.second >= 0.05f)) continue;
        const double rx = centres_l[2 * k] * s + half, ry = centres_l[2 * k + 1] * s + half;
        xs.push_back(rx);
        ys.push_back(ry);
        us.push_back(o.bx * s + half - rx);
        vs.push_back(o.by * s + half - ry);
    }
    if (xs.size() < 3) return false;
    return consensus_affine(xs, ys, us, vs, cx, cy, a, 2.0 * s);
}

int pyramid_levels(int rect_w, int rect_h) {
    const int m = std::min(rect_w, rect_h);
    int L = 0;
    while (L < 3 && (m >> (L + 1)) >= 64) ++L;
    return L;
}

void patch_centres(int rect_x, int rect_y, int rect_w, int rect_h, std::vector<int>& cxy) {
    cxy.clear();
    const int g = patch_grid(rect_w, rect_h);
    for (int j = 0; j < g; ++j)
        for (int i = 0; i < g; ++i) {
            cxy.push_back(rect_x + (int)std::lround((i + 0.5) * rect_w / g));
            cxy.push_back(rect_y + (int)std::lround((j + 0.5) * rect_h / g));
        }
}

GlobalModel fit_global_model(const std::vector<int>& cxy, const std::vector<PatchOut>& po, int rect_x,
                             int rect_y, int rect_w, int rect_h) {
    GlobalModel gm;
    gm.cx = rect_x + 0.5 * rect_w;
    gm.cy = rect_y + 0.5 * rect_h;
    std::vector<double> xs, ys, us, vs;
    for (size_t k = 0; k < po.size(); ++k) {
        const PatchOut& o = po[k];
        if (!o.ok) continue;
        if (!(o.tstd > kMinTemplateStd) || !(o.best >= kMinScore)) continue;
        double fx = 0, fy = 0;
        if (o.has_x) {
            const double den = o.xm - 2 * o.best + o.xp;
            if (den < 0) fx = std::max(-0.5, std::min(0.5, 0.5 * (o.xm - o.xp) / den));
        }
        if (o.has_y) {
            const double den = o.ym - 2 * o.best + o.yp;
            if (den < 0) fy = std::max(-0.5, std::min(0.5, 0.5 * (o.ym - o.yp) / den));
        }
        xs.push_back(cxy[2 * k]);
        ys.push_back(cxy[2 * k + 1]);
        us.push_back(o.sx + fx);
        vs.push_back(o.sy + fy);
    }
    gm.patches = (int)xs.size();
    if (gm.patches < 3) return gm;
    double a[6];
    if (!robust_affine(xs, ys, us, vs, gm.cx, gm.cy, std::max(3, (gm.patches + 1) / 2), 2.0, a) &&
        !consensus_affine(xs, ys, us, vs, gm.cx, gm.cy, a))
        return gm;
    std::copy(a, a + 6, gm.a);
    gm.locked = true;
    gm.t_u = a[0];
    gm.t_v = a[3];
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

void lattice_triangles(int LW, int LH, const std::vector<unsigned char>& kept, const std::vector<float>& rx,
                       const std::vector<float>& ry, const std::vector<float>& u, const std::vector<float>& v,
                       std::vector<TriDev>& tris, std::vector<int>& cell_tri) {
    tris.clear();
    cell_tri.assign((size_t)std::max(0, (LW - 1) * (LH - 1)) + 1, 0);
    for (int lj = 0; lj + 1 < LH; ++lj)
        for (int li = 0; li + 1 < LW; ++li) {
            const int cell = lj * (LW - 1) + li;
            cell_tri[cell] = (int)tris.size();
            const int c[4] = {lj * LW + li, lj * LW + li + 1, (lj + 1) * LW + li + 1, (lj + 1) * LW + li};
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
                TriDev t = {};
                t.cell = cell;
                double A[9], bx[3], by[3];
                for (int k = 0; k < 3; ++k) {
                    const int q = sets[s][k];
                    t.px[k] = rx[q];
                    t.py[k] = ry[q];
                    A[k * 3] = t.px[k];
                    A[k * 3 + 1] = t.py[k];
                    A[k * 3 + 2] = 1.0;
                    bx[k] = (float)(rx[q] + u[q]);
                    by[k] = (float)(ry[q] + v[q]);
                }
                double mx[3], my[3];
                if (!solve3(A, bx, mx) || !solve3(A, by, my)) continue;
                const double detF = mx[0] * my[1] - mx[1] * my[0];
                if (!finite_d(detF) || std::fabs(detF) <= 1e-3) continue;
                t.ux = mx[0] - 1.0;
                t.uy = mx[1];
                t.u = mx[2];
                t.vx = my[0];
                t.vy = my[1] - 1.0;
                t.v = my[2];
                const float x0 = std::min({t.px[0], t.px[1], t.px[2]}), x1 = std::max({t.px[0], t.px[1], t.px[2]});
                const float y0 = std::min({t.py[0], t.py[1], t.py[2]}), y1 = std::max({t.py[0], t.py[1], t.py[2]});
                t.bx0 = x0 - 15.0f;
                t.by0 = y0 - 15.0f;
                t.bx1 = x1 + 15.0f;
                t.by1 = y1 + 15.0f;
                tris.push_back(t);
            }
        }
    if (!cell_tri.empty()) cell_tri.back() = (int)tris.size();
}

}  // namespace sdic
