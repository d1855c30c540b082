// CPU ORACLE — test infrastructure only (see oracle.hpp).
//
// IncrementalSession restated on the CPU oracle's run_full_field:
//   session rules      src/pipeline/incremental_session.cpp:215-549
//   sampling, fits     src/pipeline/incremental_math.cpp:61-188
//   anchors            src/pipeline/incremental_math.cpp:190-272
//   mask warp          src/pipeline/incremental_math.cpp:274-447
//   constants          include/semper/tuning.hpp:186-281
// The distance transform is evaluated by brute force over the unsolved cells
// (exact by definition); the anchor match runs this oracle's IC-GN on the
// stored patch; the mask warp's 3x3 close treats pixels outside the image as
// neutral, as OpenCV's default morphology border does.
#include <algorithm>
#include <cmath>
#include <limits>
#include <map>

#include "oracle.hpp"

namespace orc {

namespace {

constexpr float kTaylorMaxCells = 1.5f;
constexpr int kMaskMaxSamplesPerPx = 8;
constexpr int kCoastRadiusCells = 8;
constexpr int kCoastMinNodes = 6;
constexpr float kCoastMaxTravelSubsets = 2.0f;
constexpr float kCoastTrustedTravelSubsets = 0.25f;
constexpr float kReanchorMaxShiftFraction = 0.3f;
constexpr float kReanchorMinShiftPx = 1.0f;
constexpr int kAnchorMarginPx = 10;
constexpr float kEarlyUpdateDrop = 0.05f;
constexpr float kUpdateMinReported = 0.5f;
constexpr int kMaskExtrapCells = 3;

struct Smp {
    bool ok = false, taylor = false;
    float u = 0, v = 0, corr = 0;
};

inline float node_x(const Capture& f, int i) { return (float)(f.rect_x + i * f.step); }
inline float node_y(const Capture& f, int j) { return (float)(f.rect_y + j * f.step); }

// incremental_math.cpp:61-123
Smp sample_inc(const Capture& f, float x, float y) {
    Smp s;
    if (f.gw <= 0 || f.gh <= 0 || f.step <= 0) return s;
    const float gx = (x - (float)f.rect_x) / (float)f.step;
    const float gy = (y - (float)f.rect_y) / (float)f.step;
    const int i0 = (int)std::floor(gx), j0 = (int)std::floor(gy);
    const float fx = gx - (float)i0, fy = gy - (float)j0;
    auto in = [&](int i, int j) { return i >= 0 && i < f.gw && j >= 0 && j < f.gh; };
    auto at = [&](int i, int j) { return (size_t)j * f.gw + i; };
    if (fx == 0.0f && fy == 0.0f && in(i0, j0) && f.valid[at(i0, j0)]) {
        const size_t k = at(i0, j0);
        s.ok = true;
        s.u = f.u[k];
        s.v = f.v[k];
        s.corr = f.corr[k];
        return s;
    }
    if (in(i0, j0) && in(i0 + 1, j0 + 1)) {
        const size_t a = at(i0, j0), b = at(i0 + 1, j0), c = at(i0, j0 + 1), d = at(i0 + 1, j0 + 1);
        if (f.valid[a] && f.valid[b] && f.valid[c] && f.valid[d]) {
            const float wa = (1 - fx) * (1 - fy), wb = fx * (1 - fy), wc = (1 - fx) * fy, wd = fx * fy;
            s.ok = true;
            s.u = wa * f.u[a] + wb * f.u[b] + wc * f.u[c] + wd * f.u[d];
            s.v = wa * f.v[a] + wb * f.v[b] + wc * f.v[c] + wd * f.v[d];
            s.corr = std::max(std::max(f.corr[a], f.corr[b]), std::max(f.corr[c], f.corr[d]));
            return s;
        }
    }
    const float max_d = kTaylorMaxCells * (float)f.step;
    float best = std::numeric_limits<float>::infinity();
    int bi = -1, bj = -1;
    for (int j = j0 - 1; j <= j0 + 2; ++j)
        for (int i = i0 - 1; i <= i0 + 2; ++i) {
            if (!in(i, j) || !f.valid[at(i, j)]) continue;
            const float dx = x - node_x(f, i), dy = y - node_y(f, j);
            const float d2 = dx * dx + dy * dy;
            if (d2 < best) {
                best = d2;
                bi = i;
                bj = j;
            }
        }
    if (bi < 0 || best > max_d * max_d) return s;
    const size_t k = at(bi, bj);
    const float dx = x - node_x(f, bi), dy = y - node_y(f, bj);
    s.ok = true;
    s.taylor = true;
    s.u = f.u[k] + f.ux[k] * dx + f.uy[k] * dy;
    s.v = f.v[k] + f.vx[k] * dx + f.vy[k] * dy;
    s.corr = f.corr[k];
    return s;
}

double det3(const double m[3][3]) {
    return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
           m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}

// incremental_math.cpp:125-188
Smp fit_inc(const Capture& f, float x, float y, float radius) {
    Smp s;
    if (f.gw <= 0 || f.gh <= 0 || f.step <= 0 || radius <= 0.0f) return s;
    const double st = f.step;
    const int r = (int)std::ceil(radius / st);
    const int ic = (int)std::lround((x - f.rect_x) / st), jc = (int)std::lround((y - f.rect_y) / st);
    double S[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, bu[3] = {0, 0, 0}, bv[3] = {0, 0, 0};
    int n = 0;
    float worst = 0.0f;
    for (int j = std::max(0, jc - r); j <= std::min(f.gh - 1, jc + r); ++j)
        for (int i = std::max(0, ic - r); i <= std::min(f.gw - 1, ic + r); ++i) {
            const size_t k = (size_t)j * f.gw + i;
            if (!f.valid[k]) continue;
            const double dx = node_x(f, i) - x, dy = node_y(f, j) - y;
            const double d2 = dx * dx + dy * dy;
            if (d2 > (double)radius * radius) continue;
            const double w = 1.0 / (1.0 + d2 / (st * st));
            const double pv[3] = {1.0, dx / st, dy / st};
            for (int a = 0; a < 3; ++a) {
                for (int b = 0; b < 3; ++b) S[a][b] += w * pv[a] * pv[b];
                bu[a] += w * pv[a] * f.u[k];
                bv[a] += w * pv[a] * f.v[k];
            }
            worst = std::max(worst, f.corr[k]);
            ++n;
        }
    if (n < kCoastMinNodes) return s;
    const double W = S[0][0], mx = S[0][1] / W, my = S[0][2] / W;
    const double cxx = S[1][1] / W - mx * mx, cyy = S[2][2] / W - my * my, cxy = S[1][2] / W - mx * my;
    const double tr = cxx + cyy, dt = cxx * cyy - cxy * cxy;
    if (0.5 * (tr - std::sqrt(std::max(0.0, tr * tr - 4.0 * dt))) < 1.0) return s;
    const double D = det3(S);
    if (std::abs(D) < 1e-12) return s;
    double Mu[3][3], Mv[3][3];
    for (int a = 0; a < 3; ++a) {
        Mu[a][0] = bu[a];
        Mv[a][0] = bv[a];
        for (int b = 1; b < 3; ++b) Mu[a][b] = Mv[a][b] = S[a][b];
    }
    s.ok = true;
    s.u = (float)(det3(Mu) / D);
    s.v = (float)(det3(Mv) / D);
    s.corr = worst;
    return s;
}

// distance in px from every node to the nearest "zero" position of a padded
// grid (one ring of `ring` around it), brute force
std::vector<float> dist_to_zero(const Capture& f, bool ring_is_zero) {
    std::vector<std::pair<int, int>> zeros;
    for (int j = -1; j <= f.gh; ++j)
        for (int i = -1; i <= f.gw; ++i) {
            const bool inside = i >= 0 && i < f.gw && j >= 0 && j < f.gh;
            const bool zero = inside ? !f.valid[(size_t)j * f.gw + i] : ring_is_zero;
            if (zero) zeros.push_back({i, j});
        }
    std::vector<float> d((size_t)f.gw * f.gh, std::numeric_limits<float>::infinity());
    for (int j = 0; j < f.gh; ++j)
        for (int i = 0; i < f.gw; ++i) {
            double best = std::numeric_limits<double>::infinity();
            for (const auto& z : zeros) {
                const double dx = i - z.first, dy = j - z.second;
                best = std::min(best, dx * dx + dy * dy);
            }
            d[(size_t)j * f.gw + i] = (float)std::sqrt(best) * (float)f.step;
        }
    return d;
}

enum Need { kDrop, kRefresh, kKeep };

// incremental_math.cpp:212-222
Need need_of(const Capture& f, const std::vector<float>& hole, const std::vector<float>& gap, float x, float y,
             int subset) {
    if (gap.empty()) return kKeep;
    const int i = std::min(f.gw - 1, std::max(0, (int)std::lround((x - f.rect_x) / f.step)));
    const int j = std::min(f.gh - 1, std::max(0, (int)std::lround((y - f.rect_y) / f.step)));
    if (hole[(size_t)j * f.gw + i] <= (float)(subset / 2 + f.step)) return kKeep;
    if (gap[(size_t)j * f.gw + i] <= (float)(2 * subset + 2 * f.step)) return kRefresh;
    return kDrop;
}

struct Anchor {
    std::vector<uint8_t> patch;
    int side = 0, ax = 0, ay = 0;
    float dx = 0, dy = 0;
};

// incremental_math.cpp:237-272
bool match(const Anchor& a, const Image& def, int subset, bool use6, float conv_tol, float gx, float gy,
           float max_shift, float& x, float& y, float& corr) {
    Image patch;
    image_from_u8(patch, a.side, a.side, a.patch.data());
    Subset s;
    const int c = a.side / 2;
    precompute_subset(s, patch, c, c, subset);
    if (!s.init) return false;
    s.cx = a.ax;
    s.cy = a.ay;
    Engine e;
    e.lm = true;
    e.lm_alpha = kLmAlpha;
    e.use6 = use6;
    e.conv_tol = conv_tol;
    const float g[6] = {gx - a.dx - (float)a.ax, gy - a.dy - (float)a.ay, 0, 0, 0, 0};
    const Result r = e.calculate_deformation(s, def, g, NO_SEARCH);
    if (r.status != 0 || !(r.corr <= kCorrAccept)) return false;
    const float nx = (float)a.ax + r.u + (1.0f + r.ux) * a.dx + r.uy * a.dy;
    const float ny = (float)a.ay + r.v + r.vx * a.dx + (1.0f + r.vy) * a.dy;
    if (!(std::hypot(nx - gx, ny - gy) <= max_shift)) return false;
    x = nx;
    y = ny;
    corr = r.corr;
    return true;
}

// incremental_math.cpp:274-447
std::vector<uint8_t> warp_mask(const std::vector<uint8_t>& m0, int W, int H, int gw, int gh, int rx, int ry,
                               int step, std::vector<float> U, std::vector<float> V,
                               const std::vector<uint8_t>& valid, int margin) {
    std::vector<uint8_t> out((size_t)W * H, 0);
    if (gw <= 0 || gh <= 0 || step <= 0) return out;
    // fill: empty nodes take the mean of filled 4-neighbours, pass by pass
    std::vector<uint8_t> filled(valid);
    bool any = false;
    for (size_t i = 0; i < filled.size(); ++i) {
        if (filled[i]) any = true;
        else U[i] = V[i] = 0.0f;
    }
    if (!any) return out;
    for (int pass = 0; pass < gw + gh + 2; ++pass) {
        std::vector<uint8_t> nf(filled);
        std::vector<float> nu(U), nv(V);
        bool changed = false;
        for (int y = 0; y < gh; ++y)
            for (int x = 0; x < gw; ++x) {
                const size_t i = (size_t)y * gw + x;
                if (filled[i]) continue;
                float su = 0, sv = 0;
                int cnt = 0;
                const int qx[4] = {x - 1, x + 1, x, x}, qy[4] = {y, y, y - 1, y + 1};
                for (int k = 0; k < 4; ++k) {
                    if (qx[k] < 0 || qx[k] >= gw || qy[k] < 0 || qy[k] >= gh) continue;
                    const size_t j = (size_t)qy[k] * gw + qx[k];
                    if (!filled[j]) continue;
                    su += U[j];
                    sv += V[j];
                    ++cnt;
                }
                if (!cnt) continue;
                nu[i] = su / (float)cnt;
                nv[i] = sv / (float)cnt;
                nf[i] = 1;
                changed = true;
            }
        U.swap(nu);
        V.swap(nv);
        filled.swap(nf);
        if (!changed) break;
    }
    double stretch = 1.0;
    for (int y = 0; y < gh; ++y)
        for (int x = 0; x < gw; ++x) {
            const size_t i = (size_t)y * gw + x;
            const size_t ix = x + 1 < gw ? i + 1 : i, iy = y + 1 < gh ? i + gw : i;
            const double ux = (U[ix] - U[i]) / step, uy = (U[iy] - U[i]) / step;
            const double vx = (V[ix] - V[i]) / step, vy = (V[iy] - V[i]) / step;
            stretch = std::max(stretch, std::max(std::abs(1 + ux) + std::abs(uy), std::abs(vx) + std::abs(1 + vy)));
        }
    const int per = std::min(kMaskMaxSamplesPerPx, std::max(2, (int)std::ceil(stretch / 0.7)));
    const double hs = 1.0 / per;
    const int X0 = std::max(0, rx - margin), X1 = std::min(W - 1, rx + (gw - 1) * step + margin);
    const int Y0 = std::max(0, ry - margin), Y1 = std::min(H - 1, ry + (gh - 1) * step + margin);
    auto field = [&](double X, double Y, double& u, double& v) {
        const double gxr = (X - rx) / step, gyr = (Y - ry) / step;
        const double gx = std::max(0.0, std::min(gxr, (double)(gw - 1)));
        const double gy = std::max(0.0, std::min(gyr, (double)(gh - 1)));
        const double ox = (gxr - gx) * step, oy = (gyr - gy) * step;
        int i0 = (int)std::floor(gx), j0 = (int)std::floor(gy);
        i0 = std::max(0, std::min(i0, std::max(gw - 2, 0)));
        j0 = std::max(0, std::min(j0, std::max(gh - 2, 0)));
        const int i1 = std::min(i0 + 1, gw - 1), j1 = std::min(j0 + 1, gh - 1);
        const double tx = (i1 == i0) ? 0.0 : gx - i0, ty = (j1 == j0) ? 0.0 : gy - j0;
        auto eval = [&](const std::vector<float>& F, double& val, double& dx, double& dy) {
            const double a = F[(size_t)j0 * gw + i0], b = F[(size_t)j0 * gw + i1];
            const double c = F[(size_t)j1 * gw + i0], d = F[(size_t)j1 * gw + i1];
            val = a * (1 - tx) * (1 - ty) + b * tx * (1 - ty) + c * (1 - tx) * ty + d * tx * ty;
            dx = (i1 == i0) ? 0.0 : ((b - a) * (1 - ty) + (d - c) * ty) / step;
            dy = (j1 == j0) ? 0.0 : ((c - a) * (1 - tx) + (d - b) * tx) / step;
        };
        double ux, uy, vx, vy;
        eval(U, u, ux, uy);
        eval(V, v, vx, vy);
        u += ux * ox + uy * oy;
        v += vx * ox + vy * oy;
    };
    if (X0 <= X1 && Y0 <= Y1) {
        for (int sy = Y0; sy <= Y1; ++sy)
            for (int sx = X0; sx <= X1; ++sx) {
                if (m0[(size_t)sy * W + sx] < 128) continue;
                for (int b = 0; b < per; ++b)
                    for (int a = 0; a < per; ++a) {
                        const double X = sx - 0.5 + (a + 0.5) * hs, Y = sy - 0.5 + (b + 0.5) * hs;
                        double u, v;
                        field(X, Y, u, v);
                        const long px = std::lround(X + u), py = std::lround(Y + v);
                        if (px < 0 || px >= W || py < 0 || py >= H) continue;
                        out[(size_t)py * W + px] = 255;
                    }
            }
    }
    auto morph = [&](const std::vector<uint8_t>& in, bool dil) {
        std::vector<uint8_t> r((size_t)W * H);
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                uint8_t m = dil ? 0 : 255;
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int xx = x + dx, yy = y + dy;
                        if (xx < 0 || xx >= W || yy < 0 || yy >= H) continue;
                        const uint8_t q = in[(size_t)yy * W + xx];
                        m = dil ? std::max(m, q) : std::min(m, q);
                    }
                r[(size_t)y * W + x] = m;
            }
        return r;
    };
    return morph(morph(out, true), false);
}

}  // namespace

struct IncSession {
    RefCache cache;
    RunOptions opt;
    Params p0, pk;
    int interval = 0, W = 0, H = 0, gw = 0, gh = 0;
    bool init = false;
    std::vector<uint8_t> user_mask, material, mask_k;
    int frame0_points = -1;
    std::vector<float> Ub, Vb, travel;
    std::vector<uint8_t> alive;
    std::map<size_t, Anchor> anchors;
    int ref_index = 0, since_update = 0, ref_reported = -1, adopted_reported = -1, frames = 0;
};

IncSession* inc_create() { return new IncSession(); }
void inc_destroy(IncSession* s) { delete s; }

int inc_init(IncSession* s, const uint8_t* ref, int w, int h, const uint8_t* mask, const Params& pin, int interval,
             const RunOptions& opt) {
    *s = IncSession();
    if (!ref || w <= 0 || h <= 0 || interval < 0) return kErrInit;
    Params p = pin;
    const int rc = validate_params(p, w, h);
    if (rc != 0) return rc;
    if (vsg_window_node_count(p.step, p.strain_window) < 3) return kErrStrainWindow;
    s->opt = opt;
    s->opt.capture = nullptr;
    s->p0 = p;
    s->interval = interval;
    s->W = w;
    s->H = h;
    s->gw = p.rect_w / p.step;
    s->gh = p.rect_h / p.step;
    s->material.assign((size_t)w * h, 0);
    if (mask) {
        s->user_mask.assign(mask, mask + (size_t)w * h);
        for (size_t i = 0; i < s->material.size(); ++i) s->material[i] = mask[i] >= 128 ? 255 : 0;
    } else {
        const int half = (int)std::ceil((p.subset / 2) * std::sqrt(2.0)) + 1;
        for (int y = std::max(0, p.rect_y - half); y < std::min(h, p.rect_y + p.rect_h + half); ++y)
            for (int x = std::max(0, p.rect_x - half); x < std::min(w, p.rect_x + p.rect_w + half); ++x)
                s->material[(size_t)y * w + x] = 255;
    }
    set_reference(s->cache, w, h, ref, mask);
    const size_t n = (size_t)s->gw * s->gh;
    s->Ub.assign(n, 0.0f);
    s->Vb.assign(n, 0.0f);
    s->travel.assign(n, 0.0f);
    s->alive.assign(n, 1);
    s->init = true;
    return 0;
}

int inc_process(IncSession* s, const uint8_t* def, float* out, int out_capacity, float* metrics, int metrics_len,
                IncInfo* info) {
    if (!s->init || !def || !out) return kErrInit;
    IncInfo fi;
    fi.reference_frame = s->ref_index;
    const int N = s->interval;
    bool will_update = N > 0 && s->since_update + 1 >= N;
    const size_t n0 = (size_t)s->gw * s->gh;
    const int subset = s->p0.subset;
    auto alive_count = [&]() {
        int c = 0;
        for (uint8_t a : s->alive) c += a ? 1 : 0;
        return c;
    };
    auto nothing = [&](IncInfo f) {
        s->frames++;
        f.frame0_points = std::max(s->frame0_points, 0);
        if (info) *info = f;
        const int np = metrics_len < 16 ? 0 : metrics_len >= 24 ? 24 : metrics_len == 18 ? 17 : metrics_len;
        if (metrics) std::fill_n(metrics, np, 0.0f);
        return 0;
    };
    if (s->ref_index > 0 && alive_count() == 0) return nothing(fi);

    std::vector<float> U(n0, 0.0f), V(n0, 0.0f), step_of(n0, 0.0f);
    std::vector<uint8_t> valid(n0, 0), carried(n0, 0);
    std::vector<Need> need(n0, kKeep);
    int written = 0, frame0_points = s->frame0_points;
    const float max_travel = kCoastMaxTravelSubsets * (float)subset;
    const float trusted = kCoastTrustedTravelSubsets * (float)subset;
    Capture cap;
    RunOptions ro = s->opt;
    ro.capture = &cap;
    Image defimg;
    if (s->ref_index == 0) {
        const int rc = run_full_field(s->cache, def, s->user_mask.empty() ? nullptr : s->user_mask.data(), s->p0,
                                      out, out_capacity, metrics, metrics_len, ro);
        if (rc < 0) return rc;
        written = rc;
        U = cap.u;
        V = cap.v;
        valid = cap.valid;
        carried = valid;
        if (frame0_points < 0) frame0_points = cap.attempted;
        const std::vector<float> gap = dist_to_zero(cap, true), hole = dist_to_zero(cap, false);
        for (size_t i = 0; i < n0; ++i)
            if (valid[i])
                need[i] = need_of(cap, hole, gap, node_x(cap, (int)(i % s->gw)), node_y(cap, (int)(i / s->gw)),
                                  subset);
    } else {
        const int gw = s->pk.rect_w / s->pk.step, gh = s->pk.rect_h / s->pk.step;
        std::vector<float> scratch((size_t)std::max(gw * gh, 1) * 8);
        const int rc = run_full_field(s->cache, def, s->mask_k.data(), s->pk, scratch.data(), (int)scratch.size(),
                                      metrics, metrics_len, ro);
        if (rc == kErrRoi) {
            fi.lost_this_frame = alive_count();
            std::fill(s->alive.begin(), s->alive.end(), 0);
            s->anchors.clear();
            return nothing(fi);
        }
        if (rc < 0) return rc;
        const std::vector<float> gap = dist_to_zero(cap, true), hole = dist_to_zero(cap, false);
        image_from_u8(defimg, s->W, s->H, def);
        Capture tot;
        tot.gw = s->gw;
        tot.gh = s->gh;
        tot.rect_x = s->p0.rect_x;
        tot.rect_y = s->p0.rect_y;
        tot.step = s->p0.step;
        tot.u.assign(n0, 0.0f);
        tot.v.assign(n0, 0.0f);
        tot.corr.assign(n0, kCorrInvalid);
        tot.valid.assign(n0, 0);
        for (int gy = 0; gy < s->gh; ++gy)
            for (int gx = 0; gx < s->gw; ++gx) {
                const size_t i = (size_t)gy * s->gw + gx;
                if (!s->alive[i]) continue;
                const float x0 = (float)(s->p0.rect_x + gx * s->p0.step), y0 = (float)(s->p0.rect_y + gy * s->p0.step);
                const float px = x0 + s->Ub[i], py = y0 + s->Vb[i];
                const Smp smp = sample_inc(cap, px, py);
                if (!smp.ok) {
                    // coast on the plane through the solved increments around it
                    if (px < 0.0f || py < 0.0f || px > (float)(s->W - 1) || py > (float)(s->H - 1)) continue;
                    const Smp fit = fit_inc(cap, px, py, (float)(kCoastRadiusCells * cap.step));
                    if (!fit.ok) continue;
                    const float tr = std::hypot(fit.u, fit.v);
                    if (s->travel[i] + tr > max_travel) continue;
                    U[i] = s->Ub[i] + fit.u;
                    V[i] = s->Vb[i] + fit.v;
                    carried[i] = 1;
                    step_of[i] = tr;
                    continue;
                }
                U[i] = s->Ub[i] + smp.u;
                V[i] = s->Vb[i] + smp.v;
                float corr = smp.corr;
                need[i] = need_of(cap, hole, gap, px, py, subset);
                if (s->travel[i] > trusted) {
                    const auto a = s->anchors.find(i);
                    const float max_shift = std::max(kReanchorMinShiftPx, kReanchorMaxShiftFraction * s->travel[i]);
                    float x = 0, y = 0;
                    if (need[i] == kKeep || a == s->anchors.end() ||
                        !match(a->second, defimg, subset, s->p0.use6, s->opt.conv_tol, x0 + U[i], y0 + V[i],
                               max_shift, x, y, corr)) {
                        carried[i] = 1;
                        continue;
                    }
                    U[i] = x - x0;
                    V[i] = y - y0;
                }
                valid[i] = 1;
                carried[i] = 1;
                if (smp.taylor) fi.taylor_fallbacks++;
                tot.u[i] = U[i];
                tot.v[i] = V[i];
                tot.corr[i] = corr;
                tot.valid[i] = 1;
            }
        written = pack_field(tot, s->p0.strain_window, out, out_capacity);
    }
    for (uint8_t c : carried) fi.tracked_points += c ? 1 : 0;
    fi.frame0_points = std::max(frame0_points, 0);

    int ref_reported = s->ref_reported;
    const int base = ref_reported >= 0 ? ref_reported : s->adopted_reported >= 0 ? s->adopted_reported : frame0_points;
    if (base > 0 && (float)written < kUpdateMinReported * (float)base) {
        will_update = false;
    } else if (N > 1) {
        if (ref_reported < 0) ref_reported = written;
        else if ((float)written < (1.0f - kEarlyUpdateDrop) * (float)ref_reported) will_update = true;
    }

    if (will_update) {
        const int before = (s->ref_index == 0) ? fi.frame0_points : alive_count();
        for (size_t i = 0; i < n0; ++i) {
            s->travel[i] = (carried[i] && !valid[i]) ? s->travel[i] + step_of[i] : 0.0f;
            if (!carried[i] || (valid[i] && need[i] == kDrop)) {
                s->anchors.erase(i);
            } else if (valid[i] && need[i] == kRefresh) {
                const float x = (float)(s->p0.rect_x + (int)(i % s->gw) * s->p0.step) + U[i];
                const float y = (float)(s->p0.rect_y + (int)(i / s->gw) * s->p0.step) + V[i];
                const int ax = (int)std::lround(x), ay = (int)std::lround(y), r = subset / 2 + kAnchorMarginPx;
                if (ax - r < 0 || ay - r < 0 || ax + r >= s->W || ay + r >= s->H) continue;
                Anchor an;
                an.side = 2 * r + 1;
                an.ax = ax;
                an.ay = ay;
                an.dx = x - (float)ax;
                an.dy = y - (float)ay;
                an.patch.resize((size_t)an.side * an.side);
                for (int yy = 0; yy < an.side; ++yy)
                    for (int xx = 0; xx < an.side; ++xx)
                        an.patch[(size_t)yy * an.side + xx] = def[(size_t)(ay - r + yy) * s->W + (ax - r + xx)];
                s->anchors[i] = an;
            }
        }
        const int margin = (kMaskExtrapCells + 1) * s->p0.step + subset / 2;
        std::vector<uint8_t> mk = warp_mask(s->material, s->W, s->H, s->gw, s->gh, s->p0.rect_x, s->p0.rect_y,
                                            s->p0.step, U, V, carried, margin);
        bool any_mask = false;
        for (uint8_t m : mk) any_mask = any_mask || m != 0;
        float mnx = 1e30f, mny = 1e30f, mxx = -1e30f, mxy = -1e30f;
        for (size_t i = 0; i < n0; ++i) {
            if (!carried[i]) continue;
            const float x = (float)(s->p0.rect_x + (int)(i % s->gw) * s->p0.step) + U[i];
            const float y = (float)(s->p0.rect_y + (int)(i / s->gw) * s->p0.step) + V[i];
            mnx = std::min(mnx, x);
            mxx = std::max(mxx, x);
            mny = std::min(mny, y);
            mxy = std::max(mxy, y);
        }
        Params pk = s->p0;
        pk.rect_x = pk.rect_y = pk.rect_w = pk.rect_h = 0;
        if (mnx <= mxx) {
            const int x0 = (int)std::floor(mnx), y0 = (int)std::floor(mny);
            const int xe = std::min(s->W, (int)std::ceil(mxx) + s->p0.step);
            const int ye = std::min(s->H, (int)std::ceil(mxy) + s->p0.step);
            pk.rect_x = std::max(0, x0);
            pk.rect_y = std::max(0, y0);
            pk.rect_w = std::max(0, xe - pk.rect_x);
            pk.rect_h = std::max(0, ye - pk.rect_y);
        }
        if (pk.step > 0) {
            const long long gmin = std::min(pk.rect_w / pk.step, pk.rect_h / pk.step);
            pk.strain_window = (int)std::min<long long>(pk.strain_window, 2 * (gmin + 1) * (long long)pk.step);
        }
        std::vector<uint8_t> alive(carried);
        if (!any_mask) std::fill(alive.begin(), alive.end(), 0);
        set_reference(s->cache, s->W, s->H, def, mk.data());
        s->Ub = U;
        s->Vb = V;
        s->alive = alive;
        s->mask_k = mk;
        s->pk = pk;
        s->frames++;
        s->ref_index = s->frames;
        s->since_update = 0;
        s->ref_reported = -1;
        s->adopted_reported = written;
        s->frame0_points = frame0_points;
        fi.reference_updated = 1;
        fi.lost_this_frame = std::max(0, before - alive_count());
    } else {
        s->frames++;
        s->since_update++;
        s->ref_reported = ref_reported;
        s->frame0_points = frame0_points;
    }
    if (info) *info = fi;
    return written;
}

}  // namespace orc
