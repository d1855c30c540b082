// Incremental DIC session; see incremental.hpp. Session rules:
// src/pipeline/incremental_session.cpp:215-549; sampling, coasting fits,
// anchors and the mask warp: src/pipeline/incremental_math.cpp:19-447;
// constants: include/semper/tuning.hpp:186-281.
#include "incremental.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>

#include "maskwarp.hpp"

namespace sdic {

namespace {

constexpr int kErrRoi = -2, kErrInit = -3, kErrStrainWindow = -4;
constexpr float kCorrInvalidF = -1.0f;
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

// A solved lattice as sampling wants it (NodeField, incremental_internal.hpp:18-28):
// valid = solved with corr >= 0 and kept by the displacement median test.
struct NodeField {
    int gw = 0, gh = 0, rx = 0, ry = 0, step = 0;
    const std::vector<float>*u, *v, *ux, *uy, *vx, *vy, *corr;
    const std::vector<unsigned char>* valid;
    float nx(int i) const { return (float)(rx + i * step); }
    float ny(int j) const { return (float)(ry + j * step); }
    size_t idx(int i, int j) const { return (size_t)j * gw + i; }
    bool in(int i, int j) const { return i >= 0 && i < gw && j >= 0 && j < gh; }
};

NodeField node_field(const Engine::Field& f) {
    NodeField n;
    n.gw = f.gw;
    n.gh = f.gh;
    n.rx = f.rect_x;
    n.ry = f.rect_y;
    n.step = f.step;
    n.u = &f.u;
    n.v = &f.v;
    n.ux = &f.ux;
    n.uy = &f.uy;
    n.vx = &f.vx;
    n.vy = &f.vy;
    n.corr = &f.corr;
    n.valid = &f.med;
    return n;
}

struct Sample {
    bool ok = false, taylor = false;
    float u = 0, v = 0, corr = 0;
};

double det3(const double m[3][3]) {
    return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
           m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}

// fit_increment (incremental_math.cpp:125-188): weighted plane through the
// valid nodes within radius_px, weights 1 / (1 + (d / step)^2), offsets in steps.
Sample fit_increment(const NodeField& f, float x, float y, float radius_px) {
    Sample s;
    if (f.gw <= 0 || f.gh <= 0 || f.step <= 0 || radius_px <= 0.0f) return s;
    const double step = f.step;
    const int r = (int)std::ceil(radius_px / step);
    const int ic = (int)std::lround((x - f.rx) / step), jc = (int)std::lround((y - f.ry) / step);
    const double r2max = (double)radius_px * radius_px;
    double S[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, bu[3] = {0, 0, 0}, bv[3] = {0, 0, 0};
    int n = 0;
    float worst = 0.0f;
    for (int j = std::max(0, jc - r); j <= std::min(f.gh - 1, jc + r); ++j)
        for (int i = std::max(0, ic - r); i <= std::min(f.gw - 1, ic + r); ++i) {
            const size_t k = f.idx(i, j);
            if (!(*f.valid)[k]) continue;
            const double dx = f.nx(i) - x, dy = f.ny(j) - y;
            const double d2 = dx * dx + dy * dy;
            if (d2 > r2max) continue;
            const double wgt = 1.0 / (1.0 + d2 / (step * step));
            const double p[3] = {1.0, dx / step, dy / step};
            for (int a = 0; a < 3; ++a) {
                for (int b = 0; b < 3; ++b) S[a][b] += wgt * p[a] * p[b];
                bu[a] += wgt * p[a] * (*f.u)[k];
                bv[a] += wgt * p[a] * (*f.v)[k];
            }
            worst = std::max(worst, (*f.corr)[k]);
            ++n;
        }
    if (n < kCoastMinNodes) return s;
    const double W = S[0][0];
    const double mx = S[0][1] / W, my = S[0][2] / W;
    const double cxx = S[1][1] / W - mx * mx, cyy = S[2][2] / W - my * my, cxy = S[1][2] / W - mx * my;
    const double tr = cxx + cyy, dt = cxx * cyy - cxy * cxy;
    const double lmin = 0.5 * (tr - std::sqrt(std::max(0.0, tr * tr - 4.0 * dt)));
    if (lmin < 1.0) return s;
    const double D = det3(S);
    if (std::abs(D) < 1e-12) return s;
    auto first = [&](const double b[3]) {
        double M[3][3];
        for (int a = 0; a < 3; ++a) {
            M[a][0] = b[a];
            M[a][1] = S[a][1];
            M[a][2] = S[a][2];
        }
        return det3(M) / D;
    };
    s.ok = true;
    s.u = (float)first(bu);
    s.v = (float)first(bv);
    s.corr = worst;
    return s;
}

// Exact Euclidean distance transform (to the nearest zero cell) of a w x h
// grid of 0/1, in cells: the lower envelope of parabolas per row, then per
// column (Felzenszwalb and Huttenlocher). Cells with no zero anywhere get +inf.
void edt_1d(const std::vector<double>& f, std::vector<double>& d, std::vector<int>& v, std::vector<double>& z,
            int n) {
    const double inf = std::numeric_limits<double>::infinity();
    int k = -1;
    for (int q = 0; q < n; ++q) {
        if (f[q] == inf) continue;
        if (k < 0) {
            k = 0;
            v[0] = q;
            z[0] = -inf;
            z[1] = inf;
            continue;
        }
        double s;
        while (true) {
            s = ((f[q] + (double)q * q) - (f[v[k]] + (double)v[k] * v[k])) / (2.0 * q - 2.0 * v[k]);
            if (s <= z[k] && k > 0) {
                --k;
                continue;
            }
            break;
        }
        if (s <= z[k]) {   // k == 0 and the new parabola hides it everywhere
            v[0] = q;
            z[0] = -inf;
            z[1] = inf;
            continue;
        }
        ++k;
        v[k] = q;
        z[k] = s;
        z[k + 1] = inf;
    }
    if (k < 0) {
        for (int q = 0; q < n; ++q) d[q] = inf;
        return;
    }
    int j = 0;
    for (int q = 0; q < n; ++q) {
        while (z[j + 1] < q) ++j;
        const double dq = (double)q - v[j];
        d[q] = dq * dq + f[v[j]];
    }
}

std::vector<float> edt(const std::vector<uint8_t>& one, int w, int h) {
    const double inf = std::numeric_limits<double>::infinity();
    const int m = std::max(w, h);
    std::vector<double> g((size_t)w * h), f(m), d(m), z(m + 1);
    std::vector<int> v(m);
    for (int x = 0; x < w; ++x) {
        for (int y = 0; y < h; ++y) f[y] = one[(size_t)y * w + x] ? inf : 0.0;
        edt_1d(f, d, v, z, h);
        for (int y = 0; y < h; ++y) g[(size_t)y * w + x] = d[y];
    }
    std::vector<float> out((size_t)w * h);
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) f[x] = g[(size_t)y * w + x];
        edt_1d(f, d, v, z, w);
        for (int x = 0; x < w; ++x) out[(size_t)y * w + x] = (float)std::sqrt(d[x]);
    }
    return out;
}

// gap_distance (incremental_math.cpp:190-210): px to the nearest unsolved
// node (hole; off-grid counts as solved) and to the nearest unsolved or
// off-grid position (gap).
struct GapDistance {
    std::vector<float> hole, gap;
};

GapDistance gap_distance(const NodeField& f) {
    GapDistance g;
    if (f.gw <= 0 || f.gh <= 0 || f.step <= 0) return g;
    const int W = f.gw + 2, H = f.gh + 2;
    std::vector<uint8_t> solved((size_t)W * H, 0), hole_in((size_t)W * H, 1);
    for (int j = 0; j < f.gh; ++j)
        for (int i = 0; i < f.gw; ++i) {
            const uint8_t v = (*f.valid)[f.idx(i, j)] ? 1 : 0;
            solved[(size_t)(j + 1) * W + i + 1] = v;
            hole_in[(size_t)(j + 1) * W + i + 1] = v;
        }
    const std::vector<float> dg = edt(solved, W, H), dh = edt(hole_in, W, H);
    g.gap.resize((size_t)f.gw * f.gh);
    g.hole.resize(g.gap.size());
    for (int j = 0; j < f.gh; ++j)
        for (int i = 0; i < f.gw; ++i) {
            g.gap[f.idx(i, j)] = dg[(size_t)(j + 1) * W + i + 1] * (float)f.step;
            g.hole[f.idx(i, j)] = dh[(size_t)(j + 1) * W + i + 1] * (float)f.step;
        }
    return g;
}

enum class AnchorNeed { kDrop, kRefresh, kKeep };

// anchor_need (incremental_math.cpp:212-222)
AnchorNeed anchor_need(const NodeField& f, const GapDistance& g, float x, float y, int subset) {
    if (g.gap.empty()) return AnchorNeed::kKeep;
    const int i = std::min(f.gw - 1, std::max(0, (int)std::lround((x - f.rx) / f.step)));
    const int j = std::min(f.gh - 1, std::max(0, (int)std::lround((y - f.ry) / f.step)));
    if (g.hole[f.idx(i, j)] <= (float)(subset / 2 + f.step)) return AnchorNeed::kKeep;
    if (g.gap[f.idx(i, j)] <= (float)(2 * subset + 2 * f.step)) return AnchorNeed::kRefresh;
    return AnchorNeed::kDrop;
}

// fill_field (incremental_math.cpp:278-319): every empty node with a filled
// 4-neighbour takes their mean, pass by pass from the previous pass's state.
bool fill_field(int w, int h, std::vector<float>& U, std::vector<float>& V, const std::vector<uint8_t>& valid) {
    const size_t n = (size_t)w * h;
    std::vector<uint8_t> filled(valid.begin(), valid.end());
    bool any = false;
    for (size_t i = 0; i < n; ++i) {
        if (filled[i]) any = true;
        else U[i] = V[i] = 0.0f;
    }
    if (!any) return false;
    std::vector<uint8_t> next;
    std::vector<float> nu(U), nv(V);
    for (int pass = 0; pass < w + h + 2; ++pass) {
        next = filled;
        bool changed = false;
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                const size_t i = (size_t)y * w + x;
                if (filled[i]) continue;
                float su = 0, sv = 0;
                int cnt = 0;
                const int qx[4] = {x - 1, x + 1, x, x}, qy[4] = {y, y, y - 1, y + 1};
                for (int k = 0; k < 4; ++k) {
                    if (qx[k] < 0 || qx[k] >= w || qy[k] < 0 || qy[k] >= h) continue;
                    const size_t j = (size_t)qy[k] * w + qx[k];
                    if (!filled[j]) continue;
                    su += U[j];
                    sv += V[j];
                    ++cnt;
                }
                if (cnt == 0) continue;
                nu[i] = su / (float)cnt;
                nv[i] = sv / (float)cnt;
                next[i] = 1;
                changed = true;
            }
        U = nu;
        V = nv;
        filled.swap(next);
        if (!changed) break;
    }
    return true;
}

int prefix_len(int metrics_len) {
    return metrics_len < 16 ? 0 : metrics_len >= 24 ? 24 : metrics_len == 18 ? 17 : metrics_len;
}

}  // namespace

void IncrementalSession::drop_anchors() {
    for (Anchor& a : anchors_) a.slot = -1;
    free_slots_.clear();
    slots_used_ = 0;
}

void IncrementalSession::reset() {
    eng_.reset();
    init_ = false;
    w_ = h_ = interval_ = gw_ = gh_ = 0;
    user_mask0_.clear();
    material_mask0_.clear();
    frame0_points_ = -1;
    ub_.clear();
    vb_.clear();
    ub64_.clear();
    vb64_.clear();
    c64_ = false;
    coast_travel_.clear();
    alive_.clear();
    anchors_.clear();
    free_slots_.clear();
    slots_used_ = anchor_side_ = 0;
    ref_index_ = frames_since_update_ = frame_count_ = 0;
    ref_reported_ = adopted_reported_ = -1;
    pk_ = Params{};
    mask_k_.clear();
    scratch_.clear();
}

int IncrementalSession::set_region_map(int rw, int rh, const std::vector<int>& dims) {
    if (!Engine::size_map_ok(rw, rh, dims)) return kErrInit;
    if (eng_ && eng_->set_region_map(rw, rh, dims) != 0) return kErrInit;
    region_map_ = dims;
    region_w_ = dims.empty() ? 0 : rw;
    region_h_ = dims.empty() ? 0 : rh;
    return 0;
}

int IncrementalSession::compose_on_device(std::vector<uint8_t>& flags, std::vector<double>& u64,
                                          std::vector<double>& v64, std::vector<float>& uf, std::vector<float>& vf,
                                          std::vector<float>& corr) {
    const size_t n0 = (size_t)gw_ * gh_;
    if (!c64_uf_.reserve(4 * n0) || !c64_u64_.reserve(4 * n0) || !c64_corr_.reserve(n0) ||
        !c64_flags_.reserve(2 * n0))
        return kErrInit;
    const hipStream_t st = eng_->stream();
    float* uf_d = c64_uf_.p;
    double* u64_d = c64_u64_.p;
    uint8_t* fl_d = c64_flags_.p;
    bool ok = hipMemcpyAsync(uf_d, ub_.data(), sizeof(float) * n0, hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(uf_d + n0, vb_.data(), sizeof(float) * n0, hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(fl_d, alive_.data(), n0, hipMemcpyHostToDevice, st) == hipSuccess;
    if (ok && c64_)
        ok = hipMemcpyAsync(u64_d, ub64_.data(), sizeof(double) * n0, hipMemcpyHostToDevice, st) == hipSuccess &&
             hipMemcpyAsync(u64_d + n0, vb64_.data(), sizeof(double) * n0, hipMemcpyHostToDevice, st) == hipSuccess;
    if (!ok) return kErrInit;
    ComposeTotals t{uf_d, uf_d + n0, c64_ ? u64_d : nullptr, c64_ ? u64_d + n0 : nullptr, fl_d, gw_, gh_,
                    p0_.rect_x, p0_.rect_y, p0_.step};
    ComposeOut o{c64_ ? u64_d + 2 * n0 : nullptr, c64_ ? u64_d + 3 * n0 : nullptr, uf_d + 2 * n0, uf_d + 3 * n0,
                 c64_corr_.p, fl_d + n0};
    if (launch_sample_compose64(eng_->last_dev_field, t, o, st) != hipSuccess) return kErrInit;
    flags.resize(n0);
    u64.resize(c64_ ? n0 : 0);
    v64.resize(c64_ ? n0 : 0);
    uf.resize(n0);
    vf.resize(n0);
    corr.resize(n0);
    if (c64_)
        ok = hipMemcpyAsync(u64.data(), u64_d + 2 * n0, sizeof(double) * n0, hipMemcpyDeviceToHost, st) == hipSuccess &&
             hipMemcpyAsync(v64.data(), u64_d + 3 * n0, sizeof(double) * n0, hipMemcpyDeviceToHost, st) == hipSuccess;
    ok = ok && hipMemcpyAsync(flags.data(), fl_d + n0, n0, hipMemcpyDeviceToHost, st) == hipSuccess &&
         hipMemcpyAsync(uf.data(), uf_d + 2 * n0, sizeof(float) * n0, hipMemcpyDeviceToHost, st) == hipSuccess &&
         hipMemcpyAsync(vf.data(), uf_d + 3 * n0, sizeof(float) * n0, hipMemcpyDeviceToHost, st) == hipSuccess &&
         hipMemcpyAsync(corr.data(), c64_corr_.p, sizeof(float) * n0, hipMemcpyDeviceToHost, st) == hipSuccess &&
         hipStreamSynchronize(st) == hipSuccess;
    return ok ? 0 : kErrInit;
}

int IncrementalSession::alive_count() const {
    int n = 0;
    for (uint8_t a : alive_) n += a ? 1 : 0;
    return n;
}

int IncrementalSession::nothing_to_follow(FrameInfo fi, float* metrics, int metrics_len,
                                          const std::function<void(int)>& progress, FrameInfo* info) {
    frame_count_++;
    fi.frame0_points = std::max(frame0_points_, 0);
    if (info) *info = fi;
    if (metrics) std::fill_n(metrics, prefix_len(metrics_len), 0.0f);
    if (progress) {
        try {
            progress(100);
        } catch (...) {
        }
    }
    return 0;
}

// IncrementalSession::init (incremental_session.cpp:215-265)
int IncrementalSession::init(const uint8_t* ref, int w, int h, const uint8_t* mask, const Params& pin,
                             int update_interval) {
    reset();
    if (!ref || w <= 0 || h <= 0 || update_interval < 0) return kErrInit;
    Params p = pin;
    const int rc = validate_params(p, w, h);
    if (rc != 0) return rc;
    if (vsg_window_node_count(p.step, p.strain_window) < 3) return kErrStrainWindow;
    w_ = w;
    h_ = h;
    p0_ = p;
    interval_ = update_interval;
    gw_ = p.rect_w / p.step;
    gh_ = p.rect_h / p.step;
    const size_t npx = (size_t)w * h;
    material_mask0_.assign(npx, 0);
    if (mask) {
        user_mask0_.assign(mask, mask + npx);
        for (size_t i = 0; i < npx; ++i) material_mask0_[i] = mask[i] >= 128 ? 255 : 0;
    } else {
        // the ROI rect plus the reach of its edge nodes' subsets at 45 degrees
        const int half = (int)std::ceil((p.subset / 2) * std::sqrt(2.0)) + 1;
        const int x0 = std::max(0, p.rect_x - half), y0 = std::max(0, p.rect_y - half);
        const int x1 = std::min(w, p.rect_x + p.rect_w + half), y1 = std::min(h, p.rect_y + p.rect_h + half);
        for (int y = y0; y < y1; ++y) std::memset(&material_mask0_[(size_t)y * w + x0], 255, (size_t)std::max(0, x1 - x0));
    }
    eng_ = std::make_unique<Engine>();
    eng_->use_graphs = use_graphs;
    if (!region_map_.empty() && eng_->set_region_map(region_w_, region_h_, region_map_) != 0) {
        reset();
        return kErrInit;
    }
    if (eng_->set_reference(ref, false, w, h, mask) != 0) {
        reset();
        return kErrInit;
    }
    const size_t n = (size_t)gw_ * gh_;
    ub_.assign(n, 0.0f);
    vb_.assign(n, 0.0f);
    c64_ = compose64;
    ub64_.assign(c64_ ? n : 0, 0.0);
    vb64_.assign(c64_ ? n : 0, 0.0);
    alive_.assign(n, 1);
    coast_travel_.assign(n, 0.0f);
    init_ = true;
    return 0;
}

// IncrementalSession::Impl::process (incremental_session.cpp:289-549)
namespace {
using sclk = std::chrono::steady_clock;
double ms_from(sclk::time_point t) { return std::chrono::duration<double, std::milli>(sclk::now() - t).count(); }
}  // namespace

int IncrementalSession::process(const uint8_t* def, float* out, int out_capacity, float* metrics, int metrics_len,
                                Cancel& cancel, const std::function<void(int)>& progress, FrameInfo* info) {
    const auto t_all = sclk::now();
    for (double& t : timing) t = 0.0;
    reanchor_requests = reanchor_matched = 0;
    if (!init_ || !eng_ || !out || !def) return kErrInit;
    FrameInfo fi;
    fi.reference_frame = ref_index_;
    const int N = interval_;
    bool will_update = N > 0 && frames_since_update_ + 1 >= N;
    const size_t n0 = (size_t)gw_ * gh_;
    if (ref_index_ > 0 && alive_count() == 0) return nothing_to_follow(fi, metrics, metrics_len, progress, info);

    std::vector<float> U(n0, 0.0f), V(n0, 0.0f), coast_step(n0, 0.0f);
    // the totals in double when compose64 (U, V then hold them rounded once)
    std::vector<double> U64(c64_ ? n0 : 0, 0.0), V64(c64_ ? n0 : 0, 0.0);
    auto add_increment = [&](size_t i, float du, float dv) {
        if (c64_) {
            U64[i] = ub64_[i] + (double)du;
            V64[i] = vb64_[i] + (double)dv;
            U[i] = (float)U64[i];
            V[i] = (float)V64[i];
        } else {
            U[i] = ub_[i] + du;
            V[i] = vb_[i] + dv;
        }
    };
    std::vector<uint8_t> valid(n0, 0), carried(n0, 0);
    std::vector<AnchorNeed> need(n0, AnchorNeed::kKeep);
    int coasting = 0, written = 0;
    int frame0_points = frame0_points_;
    const int subset = p0_.subset;
    const float max_travel = kCoastMaxTravelSubsets * (float)subset;
    const float trusted_travel = kCoastTrustedTravelSubsets * (float)subset;
    auto carry = [&](size_t i, float du, float dv) {
        const float travel = std::hypot(du, dv);
        if (coast_travel_[i] + travel > max_travel) return;
        add_increment(i, du, dv);
        carried[i] = 1;
        coast_step[i] = travel;
        coasting++;
    };
    auto coast = [&](size_t i, const NodeField& inc, float px, float py) {
        if (px < 0.0f || py < 0.0f || px > (float)(w_ - 1) || py > (float)(h_ - 1)) return;
        const Sample fit = fit_increment(inc, px, py, (float)(kCoastRadiusCells * inc.step));
        if (fit.ok) carry(i, fit.u, fit.v);
    };

    eng_->keep_field = true;
    if (ref_index_ == 0) {
        // the engine's own run on frame 0, straight into the caller's buffer
        const auto t_run = sclk::now();
        const int rc = eng_->run(def, false, user_mask0_.empty() ? nullptr : user_mask0_.data(), p0_, out,
                                 out_capacity, metrics, metrics_len, cancel, progress, tol);
        timing[1] = ms_from(t_run);
        if (rc < 0) return rc;
        written = rc;
        const Engine::Field& f = eng_->last_field;
        const NodeField nf = node_field(f);
        U = f.u;
        V = f.v;
        for (size_t i = 0; c64_ && i < n0; ++i) {
            U64[i] = U[i];
            V64[i] = V[i];
        }
        valid = f.med;
        carried = valid;
        if (frame0_points < 0) frame0_points = f.attempted;
        const GapDistance gd = gap_distance(nf);
        for (size_t i = 0; i < n0; ++i)
            if (valid[i])
                need[i] = anchor_need(nf, gd, nf.nx((int)(i % gw_)), nf.ny((int)(i / gw_)), subset);
    } else {
        const int gw = pk_.rect_w / pk_.step, gh = pk_.rect_h / pk_.step;
        scratch_.resize((size_t)std::max(gw * gh, 1) * 8);
        const auto t_run = sclk::now();
        const int rc = eng_->run(def, false, mask_k_.data(), pk_, scratch_.data(), (int)scratch_.size(), metrics,
                                 metrics_len, cancel, progress, tol);
        if (rc == kErrRoi) {
            // nothing on this reference's grid can be solved, now or on any later frame
            fi.lost_this_frame = alive_count();
            std::fill(alive_.begin(), alive_.end(), 0);
            drop_anchors();
            return nothing_to_follow(fi, metrics, metrics_len, progress, info);
        }
        timing[1] = ms_from(t_run);
        const auto t_comp = sclk::now();
        if (rc < 0) return rc;
        const Engine::Field inc_field = eng_->last_field;
        const NodeField inc = node_field(inc_field);
        const GapDistance gd = gap_distance(inc);
        // the sampling of every followed node's increment and its sum with
        // the node's total run on the device, in float as the reference adds
        // them or in double with FP64 composition (coasting nodes stay below)
        std::vector<uint8_t> g_flags;
        std::vector<double> g_u64, g_v64;
        std::vector<float> g_uf, g_vf, g_corr;
        {
            const int rcc = compose_on_device(g_flags, g_u64, g_v64, g_uf, g_vf, g_corr);
            if (rcc != 0) return rcc;
        }
        Engine::Field tot;
        tot.gw = gw_;
        tot.gh = gh_;
        tot.rect_x = p0_.rect_x;
        tot.rect_y = p0_.rect_y;
        tot.step = p0_.step;
        tot.u.assign(n0, 0.0f);
        tot.v.assign(n0, 0.0f);
        tot.corr.assign(n0, kCorrInvalidF);
        tot.ok.assign(n0, 0);
        // re-anchoring requests of nodes that coasted far, matched in one batch
        std::vector<size_t> ra_node;
        std::vector<uint8_t> ra_taylor;
        std::vector<Engine::AnchorMatch> ra;
        std::vector<float> corr_of(n0, 0.0f);
        for (int gy = 0; gy < gh_; ++gy)
            for (int gx = 0; gx < gw_; ++gx) {
                const size_t i = (size_t)gy * gw_ + gx;
                if (!alive_[i]) continue;
                const float x0 = (float)(p0_.rect_x + gx * p0_.step), y0 = (float)(p0_.rect_y + gy * p0_.step);
                const float px = x0 + ub_[i], py = y0 + vb_[i];
                Sample smp;
                smp.ok = (g_flags[i] & 1) != 0;
                smp.taylor = (g_flags[i] & 2) != 0;
                smp.corr = g_corr[i];
                if (!smp.ok) {
                    coast(i, inc, px, py);
                    continue;
                }
                if (c64_) {
                    U64[i] = g_u64[i];
                    V64[i] = g_v64[i];
                }
                U[i] = g_uf[i];
                V[i] = g_vf[i];
                corr_of[i] = smp.corr;
                need[i] = anchor_need(inc, gd, px, py, subset);
                if (coast_travel_[i] > trusted_travel) {
                    const Anchor* a = i < anchors_.size() && anchors_[i].slot >= 0 ? &anchors_[i] : nullptr;
                    if (need[i] == AnchorNeed::kKeep || !a) {
                        carried[i] = 1;   // unconfirmed: moves on, unreported
                        coasting++;
                        continue;
                    }
                    Engine::AnchorMatch m;
                    m.slot = a->slot;
                    m.side = anchor_side_;
                    m.ax = a->ax;
                    m.ay = a->ay;
                    m.dx = a->dx;
                    m.dy = a->dy;
                    m.guess_x = x0 + U[i];
                    m.guess_y = y0 + V[i];
                    m.max_shift = std::max(kReanchorMinShiftPx, kReanchorMaxShiftFraction * coast_travel_[i]);
                    ra_node.push_back(i);
                    ra_taylor.push_back(smp.taylor ? 1 : 0);
                    ra.push_back(m);
                    continue;
                }
                valid[i] = 1;
                carried[i] = 1;
                if (smp.taylor) fi.taylor_fallbacks++;
            }
        if (!ra.empty()) {
            if (eng_->match_anchors(ra, patch_store_.p, subset, p0_.use6, tol) != 0) return kErrInit;
            reanchor_requests = (int)ra.size();
            for (const Engine::AnchorMatch& m : ra) reanchor_matched += m.ok ? 1 : 0;
            for (size_t k = 0; k < ra.size(); ++k) {
                const size_t i = ra_node[k];
                if (!ra[k].ok) {
                    carried[i] = 1;
                    coasting++;
                    continue;
                }
                const float x0 = (float)(p0_.rect_x + (int)(i % gw_) * p0_.step);
                const float y0 = (float)(p0_.rect_y + (int)(i / gw_) * p0_.step);
                U[i] = ra[k].x - x0;
                V[i] = ra[k].y - y0;
                if (c64_) {
                    U64[i] = (double)ra[k].x - x0;
                    V64[i] = (double)ra[k].y - y0;
                }
                corr_of[i] = ra[k].corr;
                valid[i] = 1;
                carried[i] = 1;
                if (ra_taylor[k]) fi.taylor_fallbacks++;
            }
        }
        for (size_t i = 0; i < n0; ++i)
            if (valid[i]) {
                tot.u[i] = U[i];
                tot.v[i] = V[i];
                tot.corr[i] = corr_of[i];
                tot.ok[i] = 1;
            }
        // the median test runs on the total field and drops from this frame's
        // output only; it is not fed back into `valid`
        const int rcp = eng_->strain_pack(tot, p0_.strain_window, out, out_capacity, nullptr, nullptr);
        if (rcp < 0) return rcp;
        written = rcp;
        timing[2] = ms_from(t_comp);
    }

    for (uint8_t c : carried) fi.tracked_points += c ? 1 : 0;
    fi.frame0_points = std::max(frame0_points, 0);

    // the update floor and the early update on a drop in the points reported
    int ref_reported = ref_reported_;
    const int base = ref_reported >= 0 ? ref_reported : adopted_reported_ >= 0 ? adopted_reported_ : frame0_points;
    if (base > 0 && (float)written < kUpdateMinReported * (float)base) {
        will_update = false;
    } else if (N > 1) {
        if (ref_reported < 0) ref_reported = written;
        else if ((float)written < (1.0f - kEarlyUpdateDrop) * (float)ref_reported) will_update = true;
    }

    if (will_update) {
        const auto t_up = sclk::now();
        const int before = (ref_index_ == 0) ? fi.frame0_points : alive_count();
        std::vector<float> coast_travel = coast_travel_;
        for (size_t i = 0; i < n0; ++i)
            coast_travel[i] = (carried[i] && !valid[i]) ? coast_travel[i] + coast_step[i] : 0.0f;
        // the next reference's mask (warp_material_mask) and grid (prepare_reference)
        std::vector<float> Ub = U, Vb = V;
        std::vector<uint8_t> alive = carried;
        const int half = p0_.subset / 2;
        const int margin = (kMaskExtrapCells + 1) * p0_.step + half;
        std::vector<uint8_t> mask_k((size_t)w_ * h_, 0);
        {
            std::vector<float> Uf = Ub, Vf = Vb;
            const auto t_fill = sclk::now();
            const bool filled = fill_field(gw_, gh_, Uf, Vf, alive);
            timing[4] = ms_from(t_fill);
            if (filled) {
                double stretch = 1.0;
                for (int gy = 0; gy < gh_; ++gy)
                    for (int gx = 0; gx < gw_; ++gx) {
                        const size_t i = (size_t)gy * gw_ + gx;
                        const size_t ix = gx + 1 < gw_ ? i + 1 : i, iy = gy + 1 < gh_ ? i + gw_ : i;
                        const double ux = (Uf[ix] - Uf[i]) / p0_.step, uy = (Uf[iy] - Uf[i]) / p0_.step;
                        const double vx = (Vf[ix] - Vf[i]) / p0_.step, vy = (Vf[iy] - Vf[i]) / p0_.step;
                        stretch = std::max(stretch, std::max(std::abs(1 + ux) + std::abs(uy),
                                                             std::abs(vx) + std::abs(1 + vy)));
                    }
                const int per_px = std::min(kMaskMaxSamplesPerPx, std::max(2, (int)std::ceil(stretch / 0.7)));
                const int X0 = std::max(0, p0_.rect_x - margin);
                const int X1 = std::min(w_ - 1, p0_.rect_x + (gw_ - 1) * p0_.step + margin);
                const int Y0 = std::max(0, p0_.rect_y - margin);
                const int Y1 = std::min(h_ - 1, p0_.rect_y + (gh_ - 1) * p0_.step + margin);
                const MaskGrid mg{gw_, gh_, p0_.rect_x, p0_.rect_y, p0_.step};
                const auto t_warp = sclk::now();
                if (mask_warp(material_mask0_.data(), w_, h_, mg, Uf, Vf, X0, Y0, X1, Y1, per_px, mask_k, nullptr) != 0)
                    return kErrInit;
                timing[5] = ms_from(t_warp);
            }
        }
        const auto t_box = sclk::now();
        int bx0 = w_, by0 = h_, bx1 = -1, by1 = -1;
        for (int y = 0; y < h_; ++y) {
            // first and last set pixel of the row, eight bytes at a time
            const uint8_t* row = &mask_k[(size_t)y * w_];
            int first = -1, last = -1, x = 0;
            for (; x + 8 <= w_ && first < 0; x += 8) {
                uint64_t v;
                std::memcpy(&v, row + x, 8);
                if (v) first = x + __builtin_ctzll(v) / 8;
            }
            for (; x < w_ && first < 0; ++x)
                if (row[x]) first = x;
            if (first < 0) continue;
            for (x = w_; x - 8 >= first && last < 0; x -= 8) {
                uint64_t v;
                std::memcpy(&v, row + x - 8, 8);
                if (v) last = x - 8 + (63 - __builtin_clzll(v)) / 8;
            }
            for (--x; x >= first && last < 0; --x)
                if (row[x]) last = x;
            bx0 = std::min(bx0, first);
            bx1 = std::max(bx1, last);
            by0 = std::min(by0, y);
            by1 = std::max(by1, y);
        }
        Params pk = p0_;
        float mnx = 1e30f, mny = 1e30f, mxx = -1e30f, mxy = -1e30f;
        for (int gy = 0; gy < gh_; ++gy)
            for (int gx = 0; gx < gw_; ++gx) {
                const size_t i = (size_t)gy * gw_ + gx;
                if (!alive[i]) continue;
                const float x = (float)(p0_.rect_x + gx * p0_.step) + Ub[i];
                const float y = (float)(p0_.rect_y + gy * p0_.step) + Vb[i];
                mnx = std::min(mnx, x);
                mxx = std::max(mxx, x);
                mny = std::min(mny, y);
                mxy = std::max(mxy, y);
            }
        int rx = 0, ry = 0, rw = 0, rh = 0;
        if (mnx <= mxx) {
            const int x0 = (int)std::floor(mnx), y0 = (int)std::floor(mny);
            // that rect, widened by one step on the far sides, clipped to the image
            const int x1 = std::min(w_, x0 + ((int)std::ceil(mxx) - x0 + p0_.step));
            const int y1 = std::min(h_, y0 + ((int)std::ceil(mxy) - y0 + p0_.step));
            rx = std::max(0, x0);
            ry = std::max(0, y0);
            rw = std::max(0, x1 - rx);
            rh = std::max(0, y1 - ry);
        }
        pk.rect_x = rx;
        pk.rect_y = ry;
        pk.rect_w = rw;
        pk.rect_h = rh;
        // the increment solve's strain is never reported; its window only has to pass validation
        if (pk.step > 0) {
            const int64_t gmin = std::min(rw / pk.step, rh / pk.step);
            pk.strain_window = (int)std::min<int64_t>(pk.strain_window, 2 * (gmin + 1) * (int64_t)pk.step);
        }
        if (bx1 < 0) std::fill(alive.begin(), alive.end(), 0);
        // adopt the frame just solved, with the warped mask as its Ghost Wall
        timing[6] = ms_from(t_box);
        // the anchors (make_anchor, incremental_math.cpp:224-235), updated in
        // place: nothing after this point can fail and leave the session whole
        // (a failed launch below ends the session)
        if (anchors_.size() != n0) {
            drop_anchors();
            anchors_.assign(n0, Anchor{});
        }
        const int r_anchor = subset / 2 + kAnchorMarginPx;
        anchor_side_ = 2 * r_anchor + 1;
        std::vector<int32_t> cuts;   // (slot, x0, y0) per refreshed anchor
        for (size_t i = 0; i < n0; ++i) {
            Anchor& an = anchors_[i];
            if (!carried[i] || (valid[i] && need[i] == AnchorNeed::kDrop)) {
                if (an.slot >= 0) free_slots_.push_back(an.slot);
                an.slot = -1;
            } else if (valid[i] && need[i] == AnchorNeed::kRefresh) {
                // on the frame becoming the reference
                const float x = (float)(p0_.rect_x + (int)(i % gw_) * p0_.step) + U[i];
                const float y = (float)(p0_.rect_y + (int)(i / gw_) * p0_.step) + V[i];
                const int ax = (int)std::lround(x), ay = (int)std::lround(y);
                const int r = r_anchor;
                if (ax - r >= 0 && ay - r >= 0 && ax + r < w_ && ay + r < h_) {
                    if (an.slot < 0) {
                        if (!free_slots_.empty()) {
                            an.slot = free_slots_.back();
                            free_slots_.pop_back();
                        } else {
                            an.slot = slots_used_++;
                        }
                    }
                    an.ax = ax;
                    an.ay = ay;
                    an.dx = x - (float)ax;
                    an.dy = y - (float)ay;
                    cuts.push_back(an.slot);
                    cuts.push_back(ax - r);
                    cuts.push_back(ay - r);
                }
            }
        }
        if (!cuts.empty()) {
            // the store grows by doubling, keeping the slots it holds
            const size_t per = (size_t)anchor_side_ * anchor_side_;
            const size_t need_bytes = (size_t)slots_used_ * per;
            if (patch_store_.n < need_bytes) {
                DevBuf<uint8_t> grown;
                const size_t cap = std::max(need_bytes, std::max<size_t>(64 * per, 2 * patch_store_.n));
                bool ok = grown.reserve(cap);
                if (ok && patch_store_.p && patch_store_.n > 0)
                    ok = hipMemcpyAsync(grown.p, patch_store_.p, patch_store_.n, hipMemcpyDeviceToDevice,
                                        eng_->stream()) == hipSuccess;
                if (!ok) {
                    init_ = false;
                    return kErrInit;
                }
                std::swap(grown.p, patch_store_.p);
                std::swap(grown.n, patch_store_.n);
                // the old buffer goes when `grown` does, after the copy: hipFree waits for it
            }
            if (eng_->cut_patches(cuts, anchor_side_, patch_store_.p) != 0) {
                init_ = false;
                return kErrInit;
            }
        }
        const auto t_adopt = sclk::now();
        if (eng_->set_reference(def, false, w_, h_, mask_k.data()) != 0) {
            init_ = false;   // the engine lost its reference: the session cannot go on
            return kErrInit;
        }
        coast_travel_.swap(coast_travel);
        ub_.swap(Ub);
        vb_.swap(Vb);
        if (c64_) {
            ub64_.swap(U64);
            vb64_.swap(V64);
        }
        alive_.swap(alive);
        mask_k_.swap(mask_k);
        pk_ = pk;
        frame_count_++;
        ref_index_ = frame_count_;
        frames_since_update_ = 0;
        ref_reported_ = -1;
        adopted_reported_ = written;
        frame0_points_ = frame0_points;
        fi.reference_updated = true;
        fi.lost_this_frame = std::max(0, before - alive_count());
        timing[7] = ms_from(t_adopt);
        timing[3] = ms_from(t_up);
    } else {
        frame_count_++;
        frames_since_update_++;
        ref_reported_ = ref_reported;
        frame0_points_ = frame0_points;
    }
    if (info) *info = fi;
    timing[0] = ms_from(t_all);
    return written;
}

}  // namespace sdic
