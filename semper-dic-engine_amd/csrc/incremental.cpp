// Incremental DIC session; see incremental.hpp. Composition rule:
// src/pipeline/incremental_session.cpp:289-549 and incremental_math.cpp:61-123.
#include "incremental.hpp"

#include <algorithm>
#include <cmath>

namespace sdic {

namespace {
constexpr int kErrRoi = -2, kErrInit = -3, kErrStrainWindow = -4;
constexpr float kTaylorMaxCells = 1.5f;     // tuning::kIncTaylorMaxCells
constexpr float kUpdateMinReported = 0.5f;  // tuning::kIncUpdateMinReported
}  // namespace

void IncrementalSession::reset() {
    eng_.reset();
    init_ = false;
    frame_ = ref_frame_ = 0;
    alive_.clear();
    ub_.clear();
    vb_.clear();
}

int IncrementalSession::init(const uint8_t* ref, int w, int h, const uint8_t* mask, const Params& pin,
                             int update_interval) {
    reset();
    if (update_interval < 0 || !ref) return kErrInit;
    Params p = pin;
    const int rc = validate_params(p, w, h);
    if (rc != 0) return rc;
    if (vsg_window_node_count(p.step, p.strain_window) < 3) return kErrStrainWindow;
    eng_ = std::make_unique<Engine>();
    if (eng_->set_reference(ref, false, w, h, mask) != 0) {
        eng_.reset();
        return kErrInit;
    }
    w_ = w;
    h_ = h;
    p_ = p;
    interval_ = update_interval;
    gw_ = p.rect_w / p.step;
    gh_ = p.rect_h / p.step;
    mask0_ = mask ? std::vector<uint8_t>(mask, mask + (size_t)w * h) : std::vector<uint8_t>();
    // frame-0 boundary and mask scan (full_field_solver.cpp:271-317)
    alive_.assign((size_t)gw_ * gh_, 0);
    const int half = p.subset / 2, buf = half + 4 + 15;
    frame0_points_ = 0;
    for (int y = 0; y < gh_; ++y)
        for (int x = 0; x < gw_; ++x) {
            const int rx = p.rect_x + x * p.step, ry = p.rect_y + y * p.step;
            bool skip = rx - buf < 0 || rx + buf >= w || ry - buf < 0 || ry + buf >= h;
            if (!skip && mask)
                for (int dy = -half; dy <= half && !skip; ++dy)
                    for (int dx = -half; dx <= half; ++dx)
                        if (mask[(size_t)(ry + dy) * w + rx + dx] < 128) {
                            skip = true;
                            break;
                        }
            alive_[(size_t)y * gw_ + x] = skip ? 0 : 1;
            frame0_points_ += skip ? 0 : 1;
        }
    if (frame0_points_ == 0) {
        eng_.reset();
        return kErrRoi;
    }
    ub_.assign(alive_.size(), 0.0);
    vb_.assign(alive_.size(), 0.0);
    ref_first_count_ = -1;
    ref_own_count_ = frame0_points_;
    init_ = true;
    return 0;
}

int IncrementalSession::process(const uint8_t* def, float* out, int out_capacity, float* metrics, int metrics_len,
                                Cancel& cancel, const std::function<void(int)>& progress, FrameInfo* info) {
    if (!init_ || !eng_) return kErrInit;
    eng_->keep_field = true;
    const uint8_t* mask = ref_frame_ == 0 && !mask0_.empty() ? mask0_.data() : nullptr;
    const int frame = frame_ + 1;
    FrameInfo fi;
    fi.reference_frame = ref_frame_;
    fi.frame0_points = frame0_points_;
    const size_t N = alive_.size();
    std::vector<double> tu(N, 0.0), tv(N, 0.0);
    std::vector<unsigned char> meas(N, 0);
    std::vector<float> tc(N, kCorrInvalid);
    int count;
    if (ref_frame_ == 0) {
        // against the first image: the solve's own output, on the frame-0 grid
        count = eng_->run(def, false, mask, p_, out, out_capacity, metrics, metrics_len, cancel, progress);
        if (count < 0) return count;
        const Engine::Field& f = eng_->last_field;
        for (size_t i = 0; i < N; ++i)
            if (alive_[i] && f.ok[i]) {
                meas[i] = 1;
                tu[i] = f.u[i];
                tv[i] = f.v[i];
                tc[i] = f.corr[i];
            }
    } else {
        std::vector<float> scratch((size_t)std::max(8, 8 * gw_ * gh_));
        const int rc = eng_->run(def, false, mask, p_, scratch.data(), (int)scratch.size(), metrics, metrics_len,
                                 cancel, progress);
        if (rc < 0) return rc;
        const Engine::Field& f = eng_->last_field;
        auto ok = [&](int i, int j) { return i >= 0 && j >= 0 && i < f.gw && j < f.gh && f.ok[(size_t)j * f.gw + i]; };
        for (int y = 0; y < gh_; ++y)
            for (int x = 0; x < gw_; ++x) {
                const size_t n0 = (size_t)y * gw_ + x;
                if (!alive_[n0]) continue;
                // the material point in the current reference
                const double X = p_.rect_x + (double)x * p_.step + ub_[n0];
                const double Y = p_.rect_y + (double)y * p_.step + vb_[n0];
                const double fx = (X - f.rect_x) / f.step, fy = (Y - f.rect_y) / f.step;
                const int i0 = (int)std::floor(fx), j0 = (int)std::floor(fy);
                double du = 0, dv = 0;
                float corr = kCorrInvalid;
                bool have = false;
                if (ok(i0, j0) && ok(i0 + 1, j0) && ok(i0, j0 + 1) && ok(i0 + 1, j0 + 1)) {
                    const double ax = fx - i0, ay = fy - j0;
                    auto at = [&](const std::vector<float>& a, int i, int j) { return (double)a[(size_t)j * f.gw + i]; };
                    auto bil = [&](const std::vector<float>& a) {
                        return (1 - ay) * ((1 - ax) * at(a, i0, j0) + ax * at(a, i0 + 1, j0)) +
                               ay * ((1 - ax) * at(a, i0, j0 + 1) + ax * at(a, i0 + 1, j0 + 1));
                    };
                    du = bil(f.u);
                    dv = bil(f.v);
                    const int ni = std::lround(fx), nj = std::lround(fy);
                    corr = f.corr[(size_t)std::min(std::max(nj, j0), j0 + 1) * f.gw + std::min(std::max(ni, i0), i0 + 1)];
                    have = true;
                } else {
                    // first-order expansion of the nearest solved node within 1.5 cells
                    double best = 1e30;
                    int bi = -1, bj = -1;
                    for (int j = j0 - 1; j <= j0 + 2; ++j)
                        for (int i = i0 - 1; i <= i0 + 2; ++i) {
                            if (!ok(i, j)) continue;
                            const double d = std::hypot(fx - i, fy - j);
                            if (d <= kTaylorMaxCells && d < best) {
                                best = d;
                                bi = i;
                                bj = j;
                            }
                        }
                    if (bi >= 0) {
                        const size_t k = (size_t)bj * f.gw + bi;
                        const double ddx = X - (f.rect_x + (double)bi * f.step), ddy = Y - (f.rect_y + (double)bj * f.step);
                        du = f.u[k] + f.ux[k] * ddx + f.uy[k] * ddy;
                        dv = f.v[k] + f.vx[k] * ddx + f.vy[k] * ddy;
                        corr = f.corr[k];
                        have = true;
                        fi.taylor_fallbacks++;
                    }
                }
                if (!have) continue;
                meas[n0] = 1;
                tu[n0] = ub_[n0] + du;
                tv[n0] = vb_[n0] + dv;
                tc[n0] = corr;
            }
        Engine::Field tot;
        tot.gw = gw_;
        tot.gh = gh_;
        tot.rect_x = p_.rect_x;
        tot.rect_y = p_.rect_y;
        tot.step = p_.step;
        tot.u.resize(N);
        tot.v.resize(N);
        tot.corr.assign(tc.begin(), tc.end());
        tot.ok.assign(meas.begin(), meas.end());
        for (size_t i = 0; i < N; ++i) {
            tot.u[i] = (float)tu[i];
            tot.v[i] = (float)tv[i];
        }
        count = eng_->strain_pack(tot, p_.strain_window, out, out_capacity, nullptr, nullptr);
        if (count < 0) return count;
    }
    // reference update: every `interval` frames, never onto a failed solve
    bool update = interval_ > 0 && (frame - ref_frame_) % interval_ == 0;
    const int floor_base = (ref_first_count_ >= 0 && interval_ != 1) ? ref_first_count_ : ref_own_count_;
    if (update && (float)count < kUpdateMinReported * (float)floor_base) update = false;
    if (ref_first_count_ < 0) ref_first_count_ = count;
    if (update) {
        for (size_t i = 0; i < N; ++i) {
            if (!alive_[i]) continue;
            if (!meas[i]) {
                alive_[i] = 0;
                fi.lost_this_frame++;
                continue;
            }
            ub_[i] = tu[i];
            vb_[i] = tv[i];
        }
        if (eng_->set_reference(def, false, w_, h_, nullptr) != 0) return kErrInit;
        ref_frame_ = frame;
        ref_own_count_ = count;
        ref_first_count_ = -1;
    }
    frame_ = frame;
    fi.reference_updated = update;
    fi.tracked_points = (int)std::count(alive_.begin(), alive_.end(), (uint8_t)1);
    if (info) *info = fi;
    return count;
}

}  // namespace sdic
