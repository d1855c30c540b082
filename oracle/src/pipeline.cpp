// CPU ORACLE — test infrastructure only (see oracle.hpp).
// The full-field solve, restating src/pipeline/full_field_solver.cpp,
// full_field_anchors.cpp, full_field_mesh.cpp, full_field_path_{a,b,c}.cpp,
// full_field_solver_stats.cpp and reference_cache.cpp of the reference, with
// the seeding front end of seeding.hpp in place of AKAZE / Delaunay.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <queue>
#include <thread>

#include "oracle.hpp"
#include "seeding.hpp"

namespace orc {

namespace {

using clk = std::chrono::high_resolution_clock;
double ms_since(clk::time_point t0) {
    return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
}

enum Tag { kNone = 0, kInTriangle = 1, kExtrapolated = 2, kAnchor = 3, kHomography = 4 };

struct Node {  // GridPoint (full_field_internal.hpp:187-196)
    float x = 0, y = 0, u = 0, v = 0, ux = 0, uy = 0, vx = 0, vy = 0, corr = kCorrInvalid;
    bool solved = false;
    int tag = kNone;
    bool used_simplex = false;
    int iters = 0;
};

struct Stats {  // ThreadStats, one bucket per phase
    double icgn_ms = 0, simplex_ms = 0;
    int iters = 0, points = 0;
    int simplex_calls = 0, simplex_saved = 0, simplex_dead = 0;
};

void record_simplex(Stats& s, const Result& r) {
    s.simplex_calls++;
    if (r.status == 0 && r.corr <= kCorrAccept) s.simplex_saved++;
    else s.simplex_dead++;
}

// clip_axis (full_field_solver.cpp:85-95)
bool clip_axis(int& lo, int& len, int step, int extent) {
    int64_t a = lo;
    const int64_t b = std::min<int64_t>((int64_t)lo + (int64_t)len, extent);
    if (a < 0) a += ((-a + step - 1) / step) * (int64_t)step;
    if (b <= a) return false;
    lo = (int)a;
    len = (int)(b - a);
    return true;
}

int64_t max_strain_window(const Params& p) {
    if (p.step <= 0) return 0;
    const int64_t gmin = std::min(p.rect_w / p.step, p.rect_h / p.step);
    return 2 * (gmin + 1) * (int64_t)p.step;
}

// full_field_anchors.cpp:64-71
int lattice_count(int n, int stride) {
    if (n <= 0 || stride <= 0) return 0;
    int c = (n + stride - 1) / stride;
    if ((c - 1) * stride != n - 1) c++;
    return c;
}

// full_field_anchors.cpp:73-120
void plan_lattice(int gw, int gh, int& sx, int& sy) {
    int stride = (int)std::lround(std::sqrt((double)(gw * gh) / (double)kAnchorTarget));
    stride = std::max(1, stride);
    const int budget = lattice_count(gw, stride) * lattice_count(gh, stride);
    sx = stride;
    sy = stride;
    while (sx > 1 && lattice_count(gw, sx) < 3) --sx;
    while (sy > 1 && lattice_count(gh, sy) < 3) --sy;
    while (lattice_count(gw, sx) * lattice_count(gh, sy) > budget) {
        const bool gx = lattice_count(gw, sx + 1) >= 3;
        const bool gy = lattice_count(gh, sy + 1) >= 3;
        if (!gx && !gy) break;
        if (gx && (!gy || lattice_count(gw, sx) >= lattice_count(gh, sy))) ++sx;
        else ++sy;
    }
}

struct Anchor {  // AnchorResult (full_field_internal.hpp:422-433)
    int gx = 0, gy = 0;
    float rx = 0, ry = 0, u = 0, v = 0, ux = 0, uy = 0, vx = 0, vy = 0, corr = 0;
    int iters = 0;
    bool used_simplex = false, vertex = false, output = false, kept = false;
};

struct Seeds {  // MeshSeedResult
    std::vector<Anchor> anchors;
    int LW = 0, LH = 0;
    std::vector<float> ref_x, ref_y, def_x, def_y;
    int quality = 0;  // 0 NONE, 1 SPARSE, 2 FULL
    float globalU = 0, globalV = 0, coverage = 0;
    int attempted = 0, accepted = 0, solved = 0;
    bool locked = false, model = false, model_rejected = false;
    bool have_affine = false;
    double affine[6] = {};
    double acx = 0, acy = 0;
};

// full_field_anchors.cpp:122-168
void mark_kept(std::vector<Anchor>& an, int LW, int LH, const std::vector<float>& ru,
               const std::vector<float>& rv) {
    for (int lj = 0; lj < LH; ++lj)
        for (int li = 0; li < LW; ++li) {
            const size_t idx = (size_t)lj * LW + li;
            Anchor& a = an[idx];
            a.kept = false;
            if (!a.vertex) continue;
            float nu[8], nv[8];
            int nn = 0;
            for (int dj = -1; dj <= 1; ++dj)
                for (int di = -1; di <= 1; ++di) {
                    if (di == 0 && dj == 0) continue;
                    const int nj = lj + dj, ni = li + di;
                    if (ni < 0 || ni >= LW || nj < 0 || nj >= LH) continue;
                    const size_t q = (size_t)nj * LW + ni;
                    if (!an[q].vertex) continue;
                    nu[nn] = ru[q];
                    nv[nn] = rv[q];
                    ++nn;
                }
            if (nn < 3) continue;
            const float mu = median_inplace(nu, nn), mv = median_inplace(nv, nn);
            if (std::fabs(ru[idx] - mu) > kAnchorMedianTol || std::fabs(rv[idx] - mv) > kAnchorMedianTol)
                continue;
            a.kept = true;
        }
}

// Convex hull area of points (Andrew's monotone chain), for the coverage slot.
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

struct Ctx {
    const RefCache& c;
    const Params& p;
    const Image& def;
    int gw, gh;
    int threads;
    float conv_tol;
    const std::vector<PoolEntry>& pool;
    std::vector<Node>& grid;
};

Engine make_engine(const Ctx& x) {
    Engine e;
    e.lm = true;
    e.lm_alpha = kLmAlpha;
    e.use6 = x.p.use6;
    e.conv_tol = x.conv_tol;
    return e;
}

struct NodeSolve {
    bool init = false, rescue = false;
    Result r;
};

// solve_node (full_field_solver.cpp:119-139) with reject_if_ghosted.
NodeSolve solve_node(const Ctx& x, Engine& e, Subset& s, int gx, int gy, const float g[6], Mode m) {
    NodeSolve o;
    const int rx = x.p.rect_x + gx * x.p.step, ry = x.p.rect_y + gy * x.p.step;
    precompute_subset_fast(s, x.c.ref, rx, ry, x.p.subset, x.pool[(size_t)gy * x.gw + gx]);
    o.init = s.init;
    if (!o.init) return o;
    const int before = e.simplex_calls;
    o.r = e.calculate_deformation(s, x.def, g, m);
    const float gf = (float)o.r.invalid_ref / (float)(x.p.subset * x.p.subset);
    if (gf > kGhostRejectFraction) {
        o.r.corr = 2.0f;
        o.r.status = 1;
    }
    o.rescue = e.simplex_calls > before;
    return o;
}

// solve_anchor_seeds (full_field_anchors.cpp:283-574)
Seeds solve_anchor_seeds(const Ctx& x, const GlobalModel& gm, Stats& st) {
    Seeds out;
    const int gw = x.gw, gh = x.gh;
    int sx, sy;
    plan_lattice(gw, gh, sx, sy);
    std::vector<int> lx, ly;
    for (int g = 0; g < gw; g += sx) lx.push_back(g);
    if (lx.back() != gw - 1) lx.push_back(gw - 1);
    for (int g = 0; g < gh; g += sy) ly.push_back(g);
    if (ly.back() != gh - 1) ly.push_back(gh - 1);
    const int LW = (int)lx.size(), LH = (int)ly.size();
    out.locked = gm.locked;
    out.globalU = (float)gm.t_u;
    out.globalV = (float)gm.t_v;
    if (LW < 2 || LH < 2) {
        out.model = gm.full;
        return out;
    }
    const int na = LW * LH;

    auto solve_lattice = [&](bool sim, Seeds& r) -> float {
        r.locked = gm.locked;
        r.model = sim;
        r.globalU = (float)gm.t_u;
        r.globalV = (float)gm.t_v;
        r.LW = LW;
        r.LH = LH;
        std::vector<Anchor> an((size_t)na);
        const Mode mode = r.locked ? NO_SIMPLEX : AUTO_SEARCH;
        int attempted = 0;
        std::vector<Stats> ts(x.threads);
#pragma omp parallel num_threads(x.threads) reduction(+ : attempted)
        {
            Engine e = make_engine(x);
            Subset s;
            Stats& my = ts[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 4)
            for (int idx = 0; idx < na; ++idx) {
                const int gx = lx[idx % LW], gy = ly[idx / LW];
                if (x.grid[(size_t)gy * gw + gx].solved) continue;
                const int rx = x.p.rect_x + gx * x.p.step, ry = x.p.rect_y + gy * x.p.step;
                attempted++;
                float g[6] = {r.globalU, r.globalV, 0, 0, 0, 0};
                if (sim) gm.at(rx, ry, g);
                const NodeSolve ns = solve_node(x, e, s, gx, gy, g, mode);
                if (!ns.init) continue;
                my.iters += ns.r.iters;
                if (ns.rescue) record_simplex(my, ns.r);
                if (ns.r.status != 0) continue;
                Anchor& a = an[(size_t)idx];
                a.rx = (float)rx;
                a.ry = (float)ry;
                a.u = ns.r.u;
                a.v = ns.r.v;
                a.ux = ns.r.ux;
                a.uy = ns.r.uy;
                a.vx = ns.r.vx;
                a.vy = ns.r.vy;
                a.corr = ns.r.corr;
                a.gx = gx;
                a.gy = gy;
                a.iters = ns.r.iters;
                a.used_simplex = ns.rescue;
                a.vertex = ns.r.corr <= kAnchorAcceptScore;
                a.output = ns.r.corr <= kCorrAccept;
            }
            my.icgn_ms += e.t_icgn_ms;
            my.simplex_ms += e.t_simplex_ms;
        }
        for (const Stats& t : ts) {
            st.icgn_ms += t.icgn_ms;
            st.simplex_ms += t.simplex_ms;
            st.iters += t.iters;
            st.simplex_calls += t.simplex_calls;
            st.simplex_saved += t.simplex_saved;
            st.simplex_dead += t.simplex_dead;
        }
        r.attempted = attempted;
        std::vector<float> ru((size_t)na, 0.0f), rv((size_t)na, 0.0f);
        for (int i = 0; i < na; ++i) {
            const Anchor& a = an[(size_t)i];
            if (!a.vertex) continue;
            float g[6] = {0, 0, 0, 0, 0, 0};
            if (sim) gm.at(a.rx, a.ry, g);
            ru[(size_t)i] = a.u - g[0];
            rv[(size_t)i] = a.v - g[1];
        }
        mark_kept(an, LW, LH, ru, rv);
        std::vector<float> us, vs;
        std::vector<std::pair<double, double>> pts;
        for (const Anchor& a : an) {
            if (!a.kept) continue;
            r.ref_x.push_back(a.rx);
            r.ref_y.push_back(a.ry);
            r.def_x.push_back(a.rx + a.u);
            r.def_y.push_back(a.ry + a.v);
            us.push_back(a.u);
            vs.push_back(a.v);
            pts.emplace_back(a.rx, a.ry);
            if (a.output) r.solved++;
        }
        r.accepted = (int)r.ref_x.size();
        if (r.accepted >= 3) {
            r.globalU = median_inplace(us.data(), us.size());
            r.globalV = median_inplace(vs.data(), vs.size());
            const float area = (float)(x.p.rect_w * x.p.rect_h);
            r.coverage = area > 0.0f ? (float)hull_area(pts) / area : 0.0f;
        }
        const float frac = r.attempted > 0 ? (float)r.accepted / (float)r.attempted : 0.0f;
        if (r.accepted >= kAnchorFullMinVertices && frac >= kAnchorFullFraction &&
            r.coverage >= kAnchorFullCoverage)
            r.quality = 2;
        else if (r.accepted >= 3 && r.coverage >= kAnchorSparseCoverage)
            r.quality = 1;
        else
            r.quality = 0;
        r.anchors = std::move(an);
        return frac;
    };

    if (!gm.full) {
        solve_lattice(false, out);
    } else {
        const float kept = solve_lattice(true, out);
        if (out.quality != 2 || kept < kSimilarityKeepFraction) {
            out = Seeds{};
            solve_lattice(false, out);
            out.model_rejected = true;
        }
    }
    // Refit of the global model to the kept anchors (full_field_anchors.cpp:545-561):
    // a robust affine in place of the RANSAC homography.
    bool refit = false;
    if (out.accepted >= kAnchorRefitMinPoints) {
        std::vector<double> xr, yr, ud, vd;
        for (size_t i = 0; i < out.ref_x.size(); ++i) {
            xr.push_back(out.ref_x[i]);
            yr.push_back(out.ref_y[i]);
            ud.push_back((double)out.def_x[i] - out.ref_x[i]);
            vd.push_back((double)out.def_y[i] - out.ref_y[i]);
        }
        out.acx = x.p.rect_x + 0.5 * x.p.rect_w;
        out.acy = x.p.rect_y + 0.5 * x.p.rect_h;
        if (robust_affine(xr, yr, ud, vd, out.acx, out.acy, kAnchorRefitMinPoints, 1.0, out.affine))
            refit = true;
    }
    if (refit) {
        out.have_affine = true;
    } else if (gm.locked && !out.model_rejected) {
        out.have_affine = true;
        std::copy(gm.a, gm.a + 6, out.affine);
        out.acx = gm.cx;
        out.acy = gm.cy;
    }
    return out;
}

struct Guess {
    std::vector<float> u, v, ux, uy, vx, vy;
    std::vector<unsigned char> in_mesh;
};

// build_mesh_guess_field (full_field_mesh.cpp:32-282) over the lattice triangles.
Guess build_guess(const Ctx& x, const Seeds& s, double& t_mesh) {
    const auto t0 = clk::now();
    const int gw = x.gw, gh = x.gh, N = gw * gh;
    Guess g;
    g.u.assign(N, s.globalU);
    g.v.assign(N, s.globalV);
    g.ux.assign(N, 0.0f);
    g.uy.assign(N, 0.0f);
    g.vx.assign(N, 0.0f);
    g.vy.assign(N, 0.0f);
    g.in_mesh.assign(N, 0);
    std::vector<Tri> tris;
    if (s.LW >= 2 && s.LH >= 2) {
        std::vector<unsigned char> kept(s.anchors.size());
        std::vector<float> rx(kept.size()), ry(kept.size()), u(kept.size()), v(kept.size());
        for (size_t i = 0; i < kept.size(); ++i) {
            const Anchor& a = s.anchors[i];
            kept[i] = a.kept && a.rx > 0 && a.rx < x.c.w && a.ry > 0 && a.ry < x.c.h;
            rx[i] = a.rx;
            ry[i] = a.ry;
            u[i] = a.u;
            v[i] = a.v;
        }
        lattice_triangles(s.LW, s.LH, kept, rx, ry, u, v, tris);
    }
    auto assign = [&](int idx, const Tri& t, float X, float Y, int tag) {
        g.u[idx] = (float)(t.ux * X + t.uy * Y + t.u);
        g.v[idx] = (float)(t.vx * X + t.vy * Y + t.v);
        g.ux[idx] = (float)t.ux;
        g.uy[idx] = (float)t.uy;
        g.vx[idx] = (float)t.vx;
        g.vy[idx] = (float)t.vy;
        g.in_mesh[idx] = 1;
        x.grid[idx].tag = tag;
    };
    for (int y = 0; y < gh; ++y)
        for (int xx = 0; xx < gw; ++xx) {
            const int idx = y * gw + xx;
            if (x.grid[idx].solved) continue;
            const float X = (float)(x.p.rect_x + xx * x.p.step), Y = (float)(x.p.rect_y + y * x.p.step);
            for (const Tri& t : tris) {
                if (X < t.bx0 || X > t.bx1 || Y < t.by0 || Y > t.by1) continue;
                if (tri_test(t, X, Y, false) >= 0) {
                    assign(idx, t, X, Y, kInTriangle);
                    break;
                }
            }
        }
    if (s.quality == 2) {
        const float lim = -3.0f * x.p.step;
        for (int y = 0; y < gh; ++y)
            for (int xx = 0; xx < gw; ++xx) {
                const int idx = y * gw + xx;
                if (g.in_mesh[idx] || x.grid[idx].solved) continue;
                const float X = (float)(x.p.rect_x + xx * x.p.step), Y = (float)(x.p.rect_y + y * x.p.step);
                float best = lim - 1.0f;
                int bt = -1;
                for (size_t ti = 0; ti < tris.size(); ++ti) {
                    const Tri& t = tris[ti];
                    if (X < t.bx0 + lim || X > t.bx1 - lim || Y < t.by0 + lim || Y > t.by1 - lim) continue;
                    const double d = tri_test(t, X, Y, true);
                    if (d >= lim && (float)d > best) {
                        best = (float)d;
                        bt = (int)ti;
                    }
                }
                if (bt >= 0) assign(idx, tris[(size_t)bt], X, Y, kExtrapolated);
            }
    }
    if (s.have_affine) {
        for (int y = 0; y < gh; ++y)
            for (int xx = 0; xx < gw; ++xx) {
                const int idx = y * gw + xx;
                if (g.in_mesh[idx] || x.grid[idx].solved) continue;
                const double X = x.p.rect_x + (double)xx * x.p.step, Y = x.p.rect_y + (double)y * x.p.step;
                const double dx = X - s.acx, dy = Y - s.acy;
                const double* a = s.affine;
                g.u[idx] = (float)(a[0] + a[1] * dx + a[2] * dy);
                g.v[idx] = (float)(a[3] + a[4] * dx + a[5] * dy);
                g.ux[idx] = (float)a[1];
                g.uy[idx] = (float)a[2];
                g.vx[idx] = (float)a[4];
                g.vy[idx] = (float)a[5];
                g.in_mesh[idx] = 1;
                x.grid[idx].tag = kHomography;
            }
    }
    // 5x5 box smoothing with centroid correction (full_field_mesh.cpp:253-278)
    const float stepf = (float)x.p.step;
    auto smooth = [&](std::vector<float>& f, const std::vector<float>* fx, const std::vector<float>* fy) {
        const std::vector<float> tmp = f;
        for (int y = 0; y < gh; ++y)
            for (int xx = 0; xx < gw; ++xx) {
                const int idx = y * gw + xx;
                if (!g.in_mesh[idx]) continue;
                float sum = 0.0f;
                int cnt = 0, sdx = 0, sdy = 0;
                for (int dy = -2; dy <= 2; ++dy)
                    for (int dx = -2; dx <= 2; ++dx) {
                        const int ny = y + dy, nx = xx + dx;
                        if (nx >= 0 && nx < gw && ny >= 0 && ny < gh && g.in_mesh[ny * gw + nx]) {
                            sum += tmp[ny * gw + nx];
                            cnt++;
                            sdx += dx;
                            sdy += dy;
                        }
                    }
                if (cnt == 0) continue;
                f[idx] = sum / cnt;
                if (fx && (sdx != 0 || sdy != 0))
                    f[idx] -= ((*fx)[idx] * sdx + (*fy)[idx] * sdy) * stepf / cnt;
            }
    };
    smooth(g.ux, nullptr, nullptr);
    smooth(g.uy, nullptr, nullptr);
    smooth(g.vx, nullptr, nullptr);
    smooth(g.vy, nullptr, nullptr);
    smooth(g.u, &g.ux, &g.uy);
    smooth(g.v, &g.vx, &g.vy);
    t_mesh = ms_since(t0);
    return g;
}

// run_path_c (full_field_path_c.cpp:27-156)
int run_path_c(const Ctx& x, const Seeds& s, int& sx, int& sy, float& gU, float& gV, Stats& st) {
    const int gw = x.gw, gh = x.gh;
    std::vector<int> cand;
    if (!s.ref_x.empty()) {
        const float gcx = x.p.rect_x + (gw / 2.f) * x.p.step, gcy = x.p.rect_y + (gh / 2.f) * x.p.step;
        float md = 1e9f;
        int bi = -1;
        for (size_t i = 0; i < s.ref_x.size(); ++i) {
            const float dx = s.ref_x[i] - gcx, dy = s.ref_y[i] - gcy;
            const float d = dx * dx + dy * dy;
            if (d < md) {
                md = d;
                bi = (int)i;
            }
        }
        if (bi >= 0) {
            const int gx = std::max(0, std::min(gw - 1, (int)std::round((s.ref_x[bi] - x.p.rect_x) / x.p.step)));
            const int gy = std::max(0, std::min(gh - 1, (int)std::round((s.ref_y[bi] - x.p.rect_y) / x.p.step)));
            if (!x.grid[(size_t)gy * gw + gx].solved) cand.push_back(gy * gw + gx);
        }
    }
    {
        const float cx = gw / 2.0f, cy = gh / 2.0f;
        std::vector<std::pair<float, int>> fr;
        for (int y = 0; y < gh; ++y)
            for (int xx = 0; xx < gw; ++xx) {
                if (x.grid[(size_t)y * gw + xx].solved) continue;
                const float dx = xx - cx, dy = y - cy;
                fr.emplace_back(dx * dx + dy * dy, y * gw + xx);
            }
        std::stable_sort(fr.begin(), fr.end(),
                         [](const std::pair<float, int>& a, const std::pair<float, int>& b) {
                             return a.first < b.first;
                         });
        for (const auto& n : fr) {
            if ((int)cand.size() >= kPathCMaxCandidates) break;
            if (!cand.empty() && cand.front() == n.second) continue;
            cand.push_back(n.second);
        }
    }
    if (cand.empty()) return kErrRoi;
    Engine e = make_engine(x);
    Subset sub;
    const float g[6] = {gU, gV, 0, 0, 0, 0};
    int rc = kErrSeed;
    for (int flat : cand) {
        const int gx = flat % gw, gy = flat / gw;
        const NodeSolve ns = solve_node(x, e, sub, gx, gy, g, AUTO_SEARCH);
        if (!ns.init) continue;
        if (ns.r.status == 0 && ns.r.corr <= kCorrAccept) {
            sx = gx;
            sy = gy;
            gU = ns.r.u;
            gV = ns.r.v;
            rc = 0;
            break;
        }
    }
    st.icgn_ms += e.t_icgn_ms;
    st.simplex_ms += e.t_simplex_ms;
    if (rc != 0) {
        sx = cand.front() % gw;
        sy = cand.front() / gw;
    }
    return rc;
}

// run_path_a (full_field_path_a.cpp:28-114)
void run_path_a(const Ctx& x, const Guess& g, Stats& st) {
    const int N = x.gw * x.gh;
    std::vector<Stats> ts(x.threads);
#pragma omp parallel num_threads(x.threads)
    {
        Engine e = make_engine(x);
        Subset s;
        Stats& my = ts[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 32)
        for (int idx = 0; idx < N; ++idx) {
            if (!g.in_mesh[idx]) continue;
            Node& nd = x.grid[idx];
            if (nd.solved) continue;
            const int gx = idx % x.gw, gy = idx / x.gw;
            const float gs[6] = {g.u[idx], g.v[idx], g.ux[idx], g.uy[idx], g.vx[idx], g.vy[idx]};
            const bool outside = nd.tag == kExtrapolated || nd.tag == kHomography;
            const NodeSolve ns = solve_node(x, e, s, gx, gy, gs, outside ? NO_SIMPLEX : NO_SEARCH);
            if (!ns.init) continue;
            my.iters += ns.r.iters;
            if (ns.rescue) record_simplex(my, ns.r);
            if (ns.r.status == 0 && ns.r.corr <= kCorrAccept) {
                nd.u = ns.r.u;
                nd.v = ns.r.v;
                nd.ux = ns.r.ux;
                nd.uy = ns.r.uy;
                nd.vx = ns.r.vx;
                nd.vy = ns.r.vy;
                nd.corr = ns.r.corr;
                nd.solved = true;
                nd.used_simplex = ns.rescue;
                nd.iters = ns.r.iters;
                my.points++;
            } else {
                nd.solved = false;
                nd.corr = kCorrInvalid;
                nd.used_simplex = ns.rescue;
                nd.iters = ns.r.iters;
            }
        }
        my.icgn_ms += e.t_icgn_ms;
        my.simplex_ms += e.t_simplex_ms;
    }
    for (const Stats& t : ts) {
        st.icgn_ms += t.icgn_ms;
        st.simplex_ms += t.simplex_ms;
        st.iters += t.iters;
        st.points += t.points;
        st.simplex_calls += t.simplex_calls;
        st.simplex_saved += t.simplex_saved;
        st.simplex_dead += t.simplex_dead;
    }
}

struct SeedNode {
    int x, y;
    float u, v, ux, uy, vx, vy, corr;
};
// ReliabilityOrder (full_field_path_b.cpp:67-73): lower corr, then row, then column first.
struct Rel {
    bool operator()(const SeedNode& a, const SeedNode& b) const {
        if (a.corr != b.corr) return a.corr > b.corr;
        if (a.y != b.y) return a.y > b.y;
        return a.x > b.x;
    }
};
struct Cell {
    int x, y;
    float g[6];
    bool init = false, rescue = false;
    Result r;
};

// run_path_b (full_field_path_b.cpp:148-412): rounds of <= 32 popped nodes,
// claims and commits serial, two rounds in flight; the solves of the rounds
// in flight run in parallel here.
void run_path_b(const Ctx& x, const Seeds& s, float gU, float gV, int pcx, int pcy, Stats& st) {
    const int gw = x.gw, gh = x.gh;
    std::vector<unsigned char> claimed((size_t)gw * gh);
    for (int i = 0; i < gw * gh; ++i) claimed[i] = x.grid[i].solved ? 1 : 0;
    const int dx4[4] = {1, -1, 0, 0}, dy4[4] = {0, 0, 1, -1};
    std::vector<SeedNode> boundary, global;
    const SeedNode pc{pcx, pcy, gU, gV, 0, 0, 0, 0, 0};
    bool pc_queued = false;
    for (int y = 0; y < gh; ++y)
        for (int xx = 0; xx < gw; ++xx) {
            const Node& n = x.grid[(size_t)y * gw + xx];
            if (!n.solved || n.corr < 0.f) continue;
            bool touch = false;
            for (int k = 0; k < 4; ++k) {
                const int nx = xx + dx4[k], ny = y + dy4[k];
                if (nx >= 0 && nx < gw && ny >= 0 && ny < gh && !x.grid[(size_t)ny * gw + nx].solved) {
                    touch = true;
                    break;
                }
            }
            if (touch) boundary.push_back({xx, y, n.u, n.v, n.ux, n.uy, n.vx, n.vy, n.corr});
        }
    if (boundary.empty()) {
        struct C {
            int ix, iy;
            float u, v, dist, mag;
        };
        std::vector<C> cands;
        const float gcx = x.p.rect_x + (gw / 2.f) * x.p.step, gcy = x.p.rect_y + (gh / 2.f) * x.p.step;
        for (size_t i = 0; i < s.ref_x.size(); ++i) {
            const float fx = s.ref_x[i], fy = s.ref_y[i];
            const int ix = (int)std::round((fx - x.p.rect_x) / (float)x.p.step);
            const int iy = (int)std::round((fy - x.p.rect_y) / (float)x.p.step);
            if (ix < 0 || ix >= gw || iy < 0 || iy >= gh || x.grid[(size_t)iy * gw + ix].solved) continue;
            const float du = s.def_x[i] - fx, dv = s.def_y[i] - fy;
            const float wx = x.p.rect_x + ix * x.p.step, wy = x.p.rect_y + iy * x.p.step;
            const float dc = std::sqrt((wx - gcx) * (wx - gcx) + (wy - gcy) * (wy - gcy));
            cands.push_back({ix, iy, du, dv, dc, std::sqrt(du * du + dv * dv)});
        }
        if (cands.empty()) {
            global.push_back(pc);
            pc_queued = true;
        } else {
            std::sort(cands.begin(), cands.end(), [](const C& a, const C& b) {
                const float ma = std::floor(a.mag), mb = std::floor(b.mag);
                if (ma != mb) return ma < mb;
                if (a.dist != b.dist) return a.dist < b.dist;
                if (a.iy != b.iy) return a.iy < b.iy;
                return a.ix < b.ix;
            });
            for (size_t k = 0; k < cands.size() && k < 5; ++k)
                global.push_back({cands[k].ix, cands[k].iy, cands[k].u, cands[k].v, 0, 0, 0, 0, 0});
        }
    }
    std::priority_queue<SeedNode, std::vector<SeedNode>, Rel> q;
    for (const auto& b : boundary) q.push(b);

    Engine e0 = make_engine(x);
    Subset s0;
    auto solve_cell = [&](Engine& e, Subset& sub, Cell& t) {
        const NodeSolve ns = solve_node(x, e, sub, t.x, t.y, t.g, NO_SEARCH);
        t.init = ns.init;
        t.rescue = ns.rescue;
        if (t.init) t.r = ns.r;
    };
    auto commit = [&](const Cell& t) {
        Node& gp = x.grid[(size_t)t.y * gw + t.x];
        if (!t.init) return;
        st.iters += t.r.iters;
        if (t.rescue) record_simplex(st, t.r);
        if (t.r.status == 0 && t.r.corr <= kCorrAccept) {
            gp.u = t.r.u;
            gp.v = t.r.v;
            gp.ux = t.r.ux;
            gp.uy = t.r.uy;
            gp.vx = t.r.vx;
            gp.vy = t.r.vy;
            gp.corr = t.r.corr;
            gp.solved = true;
            gp.used_simplex = t.rescue;
            gp.iters = t.r.iters;
            st.points++;
            q.push({t.x, t.y, t.r.u, t.r.v, t.r.ux, t.r.uy, t.r.vx, t.r.vy, t.r.corr});
        } else {
            gp.corr = kCorrInvalid;
            gp.used_simplex = t.rescue;
            gp.iters = t.r.iters;
        }
    };
    for (const auto& sd : global) {
        const int flat = sd.y * gw + sd.x;
        if (claimed[flat]) continue;
        claimed[flat] = 1;
        Cell t{sd.x, sd.y, {sd.u, sd.v, 0, 0, 0, 0}};
        solve_cell(e0, s0, t);
        commit(t);
    }
    if (q.empty() && !global.empty() && !pc_queued) {
        const int flat = pc.y * gw + pc.x;
        if (pc.x >= 0 && pc.x < gw && pc.y >= 0 && pc.y < gh && !claimed[flat]) {
            claimed[flat] = 1;
            Cell t{pc.x, pc.y, {pc.u, pc.v, 0, 0, 0, 0}};
            solve_cell(e0, s0, t);
            commit(t);
        }
    }
    struct Round {
        std::vector<Cell> tasks;
        bool solved = false;
    };
    std::deque<Round> rounds;
    auto claim = [&]() -> bool {
        if (q.empty()) return false;
        Round r;
        int popped = 0;
        while (!q.empty() && popped < kPathBBatchNodes) {
            const SeedNode cur = q.top();
            q.pop();
            ++popped;
            for (int k = 0; k < 4; ++k) {
                const int nx = cur.x + dx4[k], ny = cur.y + dy4[k];
                if (nx < 0 || nx >= gw || ny < 0 || ny >= gh) continue;
                const int flat = ny * gw + nx;
                if (claimed[flat]) continue;
                claimed[flat] = 1;
                x.grid[flat].solved = true;
                const float ddx = (float)((nx - cur.x) * x.p.step), ddy = (float)((ny - cur.y) * x.p.step);
                Cell t{nx, ny, {cur.u + cur.ux * ddx + cur.uy * ddy, cur.v + cur.vx * ddx + cur.vy * ddy,
                                cur.ux, cur.uy, cur.vx, cur.vy}};
                r.tasks.push_back(t);
            }
        }
        rounds.push_back(std::move(r));
        return true;
    };
    std::vector<Engine> engines(x.threads, make_engine(x));
    std::vector<Subset> subs(x.threads);
    while (rounds.size() < 2 && claim()) {
    }
    while (!rounds.empty()) {
        if (!rounds.front().solved) {
            // solve every round in flight that is not solved yet
            std::vector<Cell*> todo;
            for (Round& r : rounds)
                if (!r.solved) {
                    for (Cell& c : r.tasks) todo.push_back(&c);
                    r.solved = true;
                }
            const int nt = (int)todo.size();
#pragma omp parallel for num_threads(x.threads) schedule(dynamic, 1)
            for (int i = 0; i < nt; ++i) {
                const int tid = omp_get_thread_num();
                solve_cell(engines[tid], subs[tid], *todo[i]);
            }
        }
        for (const Cell& t : rounds.front().tasks) commit(t);
        rounds.pop_front();
        while (rounds.size() < 2 && claim()) {
        }
    }
    for (const Engine& e : engines) {
        st.icgn_ms += e.t_icgn_ms;
        st.simplex_ms += e.t_simplex_ms;
    }
    st.icgn_ms += e0.t_icgn_ms;
    st.simplex_ms += e0.t_simplex_ms;
}

}  // namespace

// validate_params (full_field_solver.cpp:97-111)
int validate_params(Params& p, int width, int height) {
    if (width <= 0 || height <= 0) return kErrInit;
    if ((int64_t)width * height > (int64_t)INT32_MAX) return kErrInit;
    if (p.step <= 0) return kErrRoi;
    if (p.subset < 3 || p.subset % 2 == 0) return kErrInit;
    if (p.rect_w <= 0 || p.rect_h <= 0) return kErrRoi;
    if (!clip_axis(p.rect_x, p.rect_w, p.step, width)) return kErrRoi;
    if (!clip_axis(p.rect_y, p.rect_h, p.step, height)) return kErrRoi;
    if (p.rect_w / p.step <= 0 || p.rect_h / p.step <= 0) return kErrRoi;
    if ((int64_t)p.strain_window > max_strain_window(p)) return kErrInit;
    return 0;
}

// ReferenceCache::set_from_gray (reference_cache.cpp:50-73): gradients, then
// the Ghost Wall where the mask is below 128.
void set_reference(RefCache& c, int w, int h, const uint8_t* gray, const uint8_t* mask) {
    c = RefCache{};
    c.gray.assign(gray, gray + (size_t)w * h);
    image_from_u8(c.ref, w, h, gray);
    if (mask)
        for (size_t i = 0; i < (size_t)w * h; ++i)
            if (mask[i] < 128) c.ref.I[i] = kGhostWallValue;
    c.w = w;
    c.h = h;
    c.have = true;
}

int run_full_field(RefCache& c, const uint8_t* def_gray, const uint8_t* mask, const Params& pin,
                   float* out, int out_capacity, float* metrics, int metrics_len,
                   const RunOptions& opt) {
    const auto t_total = clk::now();
    if (out == nullptr || !c.have || def_gray == nullptr) return kErrInit;
    Params p = pin;
    const int rc0 = validate_params(p, c.w, c.h);
    if (rc0 != 0) return rc0;
    if (out_capacity < 0) out_capacity = 0;
    const int threads = opt.threads > 0 ? opt.threads : std::max(1, (int)std::thread::hardware_concurrency());

    Image def;
    image_from_u8(def, c.w, c.h, def_gray);
    const int gw = p.rect_w / p.step, gh = p.rect_h / p.step;
    if (vsg_window_node_count(p.step, p.strain_window) < 3) return kErrStrainWindow;

    std::vector<Node> grid((size_t)gw * gh);
    const bool mask_same =
        (mask == nullptr && c.grid_mask.empty()) ||
        (mask != nullptr && c.grid_mask.size() == (size_t)c.w * c.h &&
         std::equal(c.grid_mask.begin(), c.grid_mask.end(), mask));
    const bool cached = c.grid_valid && c.grid_params.rect_x == p.rect_x &&
                        c.grid_params.rect_y == p.rect_y && c.grid_params.rect_w == p.rect_w &&
                        c.grid_params.rect_h == p.rect_h && c.grid_params.step == p.step &&
                        c.grid_params.subset == p.subset && mask_same && c.gridW == gw && c.gridH == gh;
    const auto t_pre = clk::now();
    if (!cached) {
        c.grid_valid = false;
        c.skip.assign((size_t)gw * gh, 0);
        int total = 0;
        const int half = p.subset / 2, buf = half + 4 + 15;
        for (int y = 0; y < gh; ++y)
            for (int x = 0; x < gw; ++x) {
                const int rx = p.rect_x + x * p.step, ry = p.rect_y + y * p.step;
                bool skip = rx - buf < 0 || rx + buf >= c.w || ry - buf < 0 || ry + buf >= c.h;
                if (!skip && mask) {
                    for (int dy = -half; dy <= half && !skip; ++dy)
                        for (int dx = -half; dx <= half; ++dx)
                            if (mask[(size_t)(ry + dy) * c.w + rx + dx] < 128) {
                                skip = true;
                                break;
                            }
                }
                c.skip[(size_t)y * gw + x] = skip;
                if (!skip) total++;
            }
        c.total_valid = total;
        if (total == 0) return kErrRoi;
        c.pool.assign((size_t)gw * gh, PoolEntry{});
#pragma omp parallel for schedule(static) num_threads(threads)
        for (int i = 0; i < gw * gh; ++i) {
            if (c.skip[i]) continue;
            const int x = i % gw, y = i / gw;
            c.pool[i] = compute_hessian_only(c.ref, p.rect_x + x * p.step, p.rect_y + y * p.step, p.subset);
        }
        c.grid_params = p;
        c.grid_mask = mask ? std::vector<uint8_t>(mask, mask + (size_t)c.w * c.h) : std::vector<uint8_t>();
        c.gridW = gw;
        c.gridH = gh;
        c.grid_valid = true;
    }
    const double t_prepass = cached ? 0.0 : ms_since(t_pre);
    const int total_valid = c.total_valid;
    if (total_valid == 0) return kErrRoi;
    for (int y = 0; y < gh; ++y)
        for (int x = 0; x < gw; ++x) {
            Node& n = grid[(size_t)y * gw + x];
            n.x = (float)(p.rect_x + x * p.step);
            n.y = (float)(p.rect_y + y * p.step);
            n.solved = c.skip[(size_t)y * gw + x] != 0;
        }

    Ctx x{c, p, def, gw, gh, threads, opt.conv_tol, c.pool, grid};

    // seeding: global model + anchor lattice
    const auto t_seed = clk::now();
    const int L = pyramid_levels(p.rect_w, p.rect_h);
    Pyramid pr, pd;
    {
        std::vector<float> rf(c.gray.begin(), c.gray.end());
        build_pyramid(pr, rf.data(), c.w, c.h, L);
        build_pyramid(pd, def.I.data(), c.w, c.h, L);
    }
    const GlobalModel gm = global_motion(pr, pd, p.rect_x, p.rect_y, p.rect_w, p.rect_h);
    Stats st_anchor, st_a, st_b, st_c;
    Seeds seeds = solve_anchor_seeds(x, gm, st_anchor);
    const double t_seeding = ms_since(t_seed);
    int anchor_pts = 0;
    for (const Anchor& a : seeds.anchors) {
        if (!a.output || !a.kept) continue;
        Node& n = grid[(size_t)a.gy * gw + a.gx];
        n.u = a.u;
        n.v = a.v;
        n.ux = a.ux;
        n.uy = a.uy;
        n.vx = a.vx;
        n.vy = a.vy;
        n.corr = a.corr;
        n.solved = true;
        n.tag = kAnchor;
        n.used_simplex = a.used_simplex;
        n.iters = a.iters;
        anchor_pts++;
    }
    float gU = seeds.globalU, gV = seeds.globalV;
    const bool have_seed = seeds.solved > 0 || seeds.accepted >= 3;
    const bool exec_c = seeds.quality == 0 && !have_seed;
    double t_mesh = 0;
    const Guess guess = build_guess(x, seeds, t_mesh);
    int pcx = gw / 2, pcy = gh / 2;
    if (exec_c) {
        const int rc = run_path_c(x, seeds, pcx, pcy, gU, gV, st_c);
        if (rc != 0 && !seeds.locked) return rc;
    }
    run_path_a(x, guess, st_a);
    run_path_b(x, seeds, gU, gV, pcx, pcy, st_b);

    // pack_full_field_output (full_field_solver_stats.cpp:21-91)
    const auto t_strain = clk::now();
    DispField d;
    d.w = gw;
    d.h = gh;
    d.step = p.step;
    d.u.assign((size_t)gw * gh, 0.0f);
    d.v.assign((size_t)gw * gh, 0.0f);
    d.valid.assign((size_t)gw * gh, 0);
    for (int i = 0; i < gw * gh; ++i)
        if (grid[i].solved && grid[i].corr >= 0.0f) {
            d.u[i] = grid[i].u;
            d.v[i] = grid[i].v;
            d.valid[i] = 1;
        }
    const int dropped_outlier = reject_displacement_outliers(d);
    if (opt.capture) {
        Capture& cp = *opt.capture;
        const size_t n = (size_t)gw * gh;
        cp.gw = gw;
        cp.gh = gh;
        cp.rect_x = p.rect_x;
        cp.rect_y = p.rect_y;
        cp.step = p.step;
        cp.attempted = total_valid;
        cp.u.assign(n, 0.0f);
        cp.v.assign(n, 0.0f);
        cp.ux.assign(n, 0.0f);
        cp.uy.assign(n, 0.0f);
        cp.vx.assign(n, 0.0f);
        cp.vy.assign(n, 0.0f);
        cp.corr.assign(n, kCorrInvalid);
        cp.valid.assign(d.valid.begin(), d.valid.end());
        for (size_t i = 0; i < n; ++i)
            if (grid[i].solved && grid[i].corr >= 0.0f) {
                cp.u[i] = grid[i].u;
                cp.v[i] = grid[i].v;
                cp.ux[i] = grid[i].ux;
                cp.uy[i] = grid[i].uy;
                cp.vx[i] = grid[i].vx;
                cp.vy[i] = grid[i].vy;
                cp.corr[i] = grid[i].corr;
            }
    }
    const StrainField sf = compute_vsg_strain(d, p.strain_window);
    int valid = 0, dropped_post = 0;
    bool trunc = false;
    for (int y = 0; y < gh && !trunc; ++y)
        for (int xx = 0; xx < gw; ++xx) {
            const int i = y * gw + xx;
            if (!d.valid[i]) continue;
            if ((valid + 1) * 8 > out_capacity) {
                trunc = true;
                break;
            }
            if (sf.exx[i] <= kStrainFail) {
                dropped_post++;
                continue;
            }
            float* o = out + (size_t)valid * 8;
            o[0] = grid[i].x;
            o[1] = grid[i].y;
            o[2] = d.u[i];
            o[3] = d.v[i];
            o[4] = sf.exx[i];
            o[5] = sf.eyy[i];
            o[6] = sf.exy[i];
            o[7] = grid[i].corr;
            valid++;
        }
    const double t_str = ms_since(t_strain);
    const double total_ms = ms_since(t_total);

    // fill_engine_metrics (full_field_solver_stats.cpp:182-280)
    const int ncopy = metrics_len < 16 ? 0 : metrics_len >= 24 ? 24 : metrics_len == 18 ? 17 : metrics_len;
    if (metrics && ncopy > 0) {
        const int iters = st_a.iters + st_b.iters + st_anchor.iters;
        float m[24];
        m[0] = (float)total_valid;
        m[1] = (float)valid;
        m[2] = (float)(total_valid - valid);
        m[3] = (float)st_a.points;
        m[4] = (float)st_b.points;
        m[5] = (float)(st_a.simplex_calls + st_b.simplex_calls + st_anchor.simplex_calls);
        m[6] = (float)(st_a.simplex_saved + st_b.simplex_saved + st_anchor.simplex_saved);
        m[7] = (float)(st_a.simplex_dead + st_b.simplex_dead + st_anchor.simplex_dead);
        m[8] = valid > 0 ? (float)iters / (float)valid : 0.0f;
        m[9] = (float)total_ms;
        m[10] = (float)t_seeding;
        m[11] = (float)t_prepass;
        m[12] = (float)t_mesh;
        m[13] = (float)t_str;
        m[14] = total_ms > 0 ? (float)(valid / total_ms) : 0.0f;
        m[15] = total_valid > 0 ? ((float)valid / total_valid) * 100.0f : 0.0f;
        m[16] = (float)seeds.quality;
        m[17] = (float)(st_a.simplex_ms + st_b.simplex_ms + st_anchor.simplex_ms + st_c.simplex_ms);
        m[18] = (float)(st_a.icgn_ms + st_b.icgn_ms + st_anchor.icgn_ms + st_c.icgn_ms);
        m[19] = (float)anchor_pts;
        m[20] = (float)dropped_post;
        m[21] = seeds.locked ? 1.0f : 0.0f;
        m[22] = seeds.coverage;
        m[23] = (float)dropped_outlier;
        for (int i = 0; i < ncopy; ++i) metrics[i] = m[i];
    }
    return valid;
}

int pack_field(const Capture& f, int window, float* out, int out_capacity) {
    DispField d;
    d.w = f.gw;
    d.h = f.gh;
    d.step = f.step;
    d.u = f.u;
    d.v = f.v;
    d.valid = f.valid;
    reject_displacement_outliers(d);
    const StrainField sf = compute_vsg_strain(d, window);
    int valid = 0;
    for (int y = 0; y < f.gh; ++y)
        for (int x = 0; x < f.gw; ++x) {
            const int i = y * f.gw + x;
            if (!d.valid[i]) continue;
            if ((valid + 1) * 8 > out_capacity) return valid;
            if (sf.exx[i] <= kStrainFail) continue;
            float* o = out + (size_t)valid * 8;
            o[0] = (float)(f.rect_x + x * f.step);
            o[1] = (float)(f.rect_y + y * f.step);
            o[2] = d.u[i];
            o[3] = d.v[i];
            o[4] = sf.exx[i];
            o[5] = sf.eyy[i];
            o[6] = sf.exy[i];
            o[7] = f.corr[i];
            valid++;
        }
    return valid;
}

}  // namespace orc
