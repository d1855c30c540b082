// CPU ORACLE — test infrastructure only (see oracle.hpp).
// IC-GN, Nelder-Mead rescue and coarse search, restating
// src/math/optimization_engine.cpp and include/semper/simd.hpp of the reference.
#include <algorithm>
#include <chrono>

#include "oracle.hpp"

namespace orc {

namespace {

using clk = std::chrono::high_resolution_clock;

// optimization_engine.cpp:19-21 — summed in double, rounded to float once.
inline float warped(float c, float a, float b, float t, float x, float y) {
    return (float)((double)c + ((double)a * x + (double)b * y + (double)t));
}

// optimization_engine.cpp:28-32 — the rounded pixel must be 4 px inside.
inline bool inside_guard(float x, float y, int w, int h) {
    const float rx = x + 0.5f, ry = y + 0.5f;
    return rx >= 4.0f && rx < (float)(w - 4) && ry >= 4.0f && ry < (float)(h - 4);
}

// 3x3 affine warp as a full 3x3 matrix, row-major.
struct M3 {
    float m[9];
};

M3 warp_matrix(const float p[6]) {
    // p = u, v, ux, uy, vx, vy (optimization_engine.cpp:34-43)
    return M3{{1.0f + p[2], p[3], p[0], p[4], 1.0f + p[5], p[1], 0.0f, 0.0f, 1.0f}};
}

// Cofactor inverse, det from the first column (Eigen's 3x3 closed form).
M3 inverse3(const M3& a) {
    auto A = [&](int r, int c) { return a.m[r * 3 + c]; };
    auto cof = [&](int i, int j) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        return A(i1, j1) * A(i2, j2) - A(i1, j2) * A(i2, j1);
    };
    const float c00 = cof(0, 0), c10 = cof(1, 0), c20 = cof(2, 0);
    const float det = c00 * A(0, 0) + c10 * A(1, 0) + c20 * A(2, 0);
    const float inv = 1.0f / det;
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i * 3 + j] = cof(j, i) * inv;
    return r;
}

M3 mul3(const M3& a, const M3& b) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            float s = 0.0f;
            for (int k = 0; k < 3; ++k) s += a.m[i * 3 + k] * b.m[k * 3 + j];
            r.m[i * 3 + j] = s;
        }
    return r;
}

Result result_from(const M3& W, int status, float score, int iters, int invalid_ref) {
    Result r;
    r.u = W.m[2];
    r.v = W.m[5];
    r.ux = W.m[0] - 1.0f;
    r.uy = W.m[1];
    r.vx = W.m[3];
    r.vy = W.m[4] - 1.0f;
    r.status = status;
    r.corr = score;
    r.iters = iters;
    r.invalid_ref = invalid_ref;
    return r;
}

// simd.hpp's 128-bit kernels keep four partial sums (lane i % 4) and fold
// them as (a0 + a2) + (a1 + a3), then add the scalar tail in order.
struct Lanes4 {
    float a[4] = {0, 0, 0, 0};
    float fold() const { return (a[0] + a[2]) + (a[1] + a[3]); }
};

float sum_sq_diff(const float* v, int n, float mean) {
    Lanes4 acc;
    int i = 0;
    for (; i + 4 <= n; i += 4)
        for (int k = 0; k < 4; ++k) {
            const float d = v[i + k] - mean;
            acc.a[k] = d * d + acc.a[k];
        }
    float s = acc.fold();
    for (; i < n; ++i) {
        const float d = v[i] - mean;
        s += d * d;
    }
    return s;
}

float znssd_sum(const float* v, const float* ref, int n, float mean, float inv) {
    Lanes4 acc;
    int i = 0;
    for (; i + 4 <= n; i += 4)
        for (int k = 0; k < 4; ++k) {
            const float nd = (v[i + k] - mean) * inv;
            const float d = ref[i + k] - nd;
            acc.a[k] = d * d + acc.a[k];
        }
    float s = acc.fold();
    for (; i < n; ++i) {
        const float nd = (v[i] - mean) * inv;
        const float d = ref[i] - nd;
        s += d * d;
    }
    return s;
}

float znssd_error_and_gradient(const float* v, const float* ref, const float* sd, int n,
                               float mean, float inv, float dp[6]) {
    Lanes4 e, g[6];
    int i = 0;
    for (; i + 4 <= n; i += 4)
        for (int k = 0; k < 4; ++k) {
            const float nd = (v[i + k] - mean) * inv;
            const float d = ref[i + k] - nd;
            e.a[k] = d * d + e.a[k];
            for (int q = 0; q < 6; ++q) g[q].a[k] = sd[q * n + i + k] * d + g[q].a[k];
        }
    float err = e.fold();
    for (int q = 0; q < 6; ++q) dp[q] = g[q].fold();
    for (; i < n; ++i) {
        const float nd = (v[i] - mean) * inv;
        const float d = ref[i] - nd;
        err += d * d;
        for (int q = 0; q < 6; ++q) dp[q] += sd[q * n + i] * d;
    }
    return err;
}

}  // namespace

// optimization_engine.cpp:191-269 (batched and single calls give the same value).
static int sample_w(const Engine& e, const Subset& s, const Image& def, const M3& W,
                    std::vector<float>& vals, std::vector<unsigned char>& valid, float& sum) {
    const int n = s.n;
    int count = 0;
    sum = 0.0f;
    for (int i = 0; i < n; ++i) {
        const float x = s.ox[i], y = s.oy[i];
        const float fx = warped((float)s.cx, W.m[0], W.m[1], W.m[2], x, y);
        const float fy = warped((float)s.cy, W.m[3], W.m[4], W.m[5], x, y);
        if (!inside_guard(fx, fy, def.w, def.h)) {
            vals[i] = 0.0f;
            valid[i] = 0;
            continue;
        }
        const float v = e.interp(def, fx, fy);
        vals[i] = v;
        valid[i] = (s.ref[i] < kGhostWallThreshold) ? 0 : 1;
        if (valid[i]) {
            sum += v;
            count++;
        }
    }
    return count;
}

// optimization_engine.cpp:55-118.
Result Engine::calculate_deformation(const Subset& s, const Image& def, const float g[6],
                                     Mode mode) {
    if ((int)buf_icgn.size() != s.n) buf_icgn.resize(s.n);
    if ((int)buf_simplex.size() != s.n) buf_simplex.resize(s.n);
    val_icgn.resize(s.n);
    val_simplex.resize(s.n);

    auto run_icgn = [&](const float p[6]) {
        const auto t0 = clk::now();
        Result r = solve_icgn(s, def, p);
        t_icgn_ms += std::chrono::duration<double, std::milli>(clk::now() - t0).count();
        return r;
    };
    auto run_simplex = [&](const Result& st, bool trans) {
        const auto t0 = clk::now();
        Result r = solve_simplex(s, def, st, trans);
        t_simplex_ms += std::chrono::duration<double, std::milli>(clk::now() - t0).count();
        simplex_calls++;
        return r;
    };

    Result res;
    if (mode == AUTO_SEARCH) {
        float eu = g[0], ev = g[1];
        if (eu == 0.0f && ev == 0.0f) estimate_initial_guess(s, def, eu, ev);
        Result st;
        st.u = eu;
        st.v = ev;
        st.corr = 1.0f;
        const Result coarse = run_simplex(st, true);
        const float p[6] = {coarse.u, coarse.v, 0.0f, 0.0f, 0.0f, 0.0f};
        res = run_icgn(p);
        res.trigger_iters = res.iters;
    } else {
        res = run_icgn(g);
        res.trigger_iters = res.iters;
        if ((res.status != 0 || res.corr > kCorrSimplexTrigger) && mode != NO_SIMPLEX) {
            Result st;
            st.u = g[0];
            st.v = g[1];
            st.ux = g[2];
            st.uy = g[3];
            st.vx = g[4];
            st.vy = g[5];
            st.corr = 1.0f;
            const Result rescue = run_simplex(st, false);
            const int trig = res.iters;
            const float p[6] = {rescue.u, rescue.v, rescue.ux, rescue.uy, rescue.vx, rescue.vy};
            Result rescued = run_icgn(p);
            rescued.trigger_iters = trig;
            const bool first_better =
                res.status == 0 && (rescued.status != 0 || res.corr < rescued.corr);
            if (!first_better) res = rescued;
        }
    }
    return res;
}

// optimization_engine.cpp:120-179.
void Engine::estimate_initial_guess(const Subset& s, const Image& def, float& bu, float& bv) {
    const int n = s.n;
    std::vector<float> r, d;
    r.reserve(n / 9 + 1);
    d.reserve(n / 9 + 1);
    int sampled = 0;
    for (int i = 0; i < n; i += 9) sampled++;
    float best = 1e20f;
    for (int v = -15; v <= 15; v += 2) {
        for (int u = -15; u <= 15; u += 2) {
            r.clear();
            d.clear();
            for (int i = 0; i < n; i += 9) {
                if (s.ref[i] < kGhostWallThreshold) continue;
                const float x = s.cx + s.ox[i] + (float)u;
                const float y = s.cy + s.oy[i] + (float)v;
                if (!inside_guard(x, y, def.w, def.h)) continue;
                r.push_back(s.ref[i]);
                d.push_back(interp(def, x, y));
            }
            if (r.size() < 3 || 2 * (int)r.size() < sampled) continue;
            const float m = (float)r.size();
            float mr = 0.0f, md = 0.0f;
            for (size_t k = 0; k < r.size(); ++k) {
                mr += r[k];
                md += d[k];
            }
            mr /= m;
            md /= m;
            float vr = 0.0f, vd = 0.0f;
            for (size_t k = 0; k < r.size(); ++k) {
                vr += (r[k] - mr) * (r[k] - mr);
                vd += (d[k] - md) * (d[k] - md);
            }
            const float sr = std::sqrt(vr / m), sdv = std::sqrt(vd / m);
            if (!(sr > 1e-5f) || !(sdv > 1e-5f)) continue;
            float score = 0.0f;
            for (size_t k = 0; k < r.size(); ++k) {
                const float e = (r[k] - mr) / sr - (d[k] - md) / sdv;
                score += e * e;
            }
            score /= m;
            if (score < best) {
                best = score;
                bu = (float)u;
                bv = (float)v;
            }
        }
    }
}

// optimization_engine.cpp:272-465.
Result Engine::solve_icgn(const Subset& s, const Image& def, const float g[6]) {
    const int n = s.n, half = s.dim / 2;
    M3 W = warp_matrix(g);
    float Hs[36];
    std::memcpy(Hs, s.Hinv, sizeof Hs);
    if (lm && lm_alpha > 0.0f) {
        float Hd[36];
        std::memcpy(Hd, s.H, sizeof Hd);
        Hd[0] += lm_alpha;
        Hd[7] += lm_alpha;
        if (well_conditioned(Hd, half)) invert6(Hd, Hs);
    }
    std::vector<float>& vals = buf_icgn;
    std::vector<unsigned char>& valid = val_icgn;
    float final_score = 1.0f;
    int invalid_ref = 0;
    for (int k = 0; k < n; ++k)
        if (s.ref[k] < kGhostWallThreshold) invalid_ref++;
    auto reject = [&](int it) { return result_from(W, 1, 2.0f, it, invalid_ref); };
    const int min_valid = (int)((float)n * kMinValidPixelFraction);

    for (int iter = 0; iter < kIcgnMaxIter; ++iter) {
        float dsum = 0.0f;
        const int nv = sample_w(*this, s, def, W, vals, valid, dsum);
        if (nv < min_valid) return reject(iter);
        const float dmean = dsum / (float)nv;
        float err = 0.0f;
        float dp[6] = {0, 0, 0, 0, 0, 0};
        float Hm[36];
        const float* Hstep = Hs;
        if (nv == n) {
            const float ssq = sum_sq_diff(vals.data(), n, dmean);
            float dstd = std::sqrt(ssq / nv);
            if (dstd < 1e-5f) dstd = 1.0f;
            const float inv = 1.0f / dstd;
            err = znssd_error_and_gradient(vals.data(), s.nref.data(), s.sd.data(), n, dmean, inv,
                                           dp);
        } else {
            float Hdyn[36] = {};
            float ssq = 0.0f;
            for (int i = 0; i < n; ++i)
                if (valid[i]) {
                    const float d = vals[i] - dmean;
                    ssq += d * d;
                }
            float dstd = std::sqrt(ssq / nv);
            if (dstd < 1e-5f) dstd = 1.0f;
            const float inv = 1.0f / dstd;
            for (int i = 0; i < n; ++i) {
                if (!valid[i]) continue;
                const float nd = (vals[i] - dmean) * inv;
                const float d = s.nref[i] - nd;
                err += d * d;
                float sdi[6];
                for (int q = 0; q < 6; ++q) {
                    sdi[q] = s.sd[q * n + i];
                    dp[q] += sdi[q] * d;
                }
                for (int r = 0; r < 6; ++r)
                    for (int c = 0; c < 6; ++c) Hdyn[r * 6 + c] += sdi[r] * sdi[c];
            }
            const float det2 = Hdyn[0] * Hdyn[7] - Hdyn[6] * Hdyn[1];
            const float nrm2 = Hdyn[0] * Hdyn[0] + Hdyn[1] * Hdyn[1] + Hdyn[6] * Hdyn[6] +
                               Hdyn[7] * Hdyn[7];
            const float cond2 = (std::fabs(det2) > 1e-12f) ? (nrm2 / std::fabs(det2)) : 1.0e13f;
            if (cond2 > 1.0e12f) return reject(iter);
            if (lm && lm_alpha > 0.0f) {
                Hdyn[0] += lm_alpha;
                Hdyn[7] += lm_alpha;
            }
            if (!well_conditioned(Hdyn, half)) return reject(iter);
            invert6(Hdyn, Hm);
            Hstep = Hm;
        }
        final_score = err / (float)nv;

        float d[6];
        for (int r = 0; r < 6; ++r) {
            float acc = 0.0f;
            for (int c = 0; c < 6; ++c) acc += Hstep[r * 6 + c] * dp[c];
            d[r] = -acc;
        }
        M3 dW{{1.0f + d[2], d[3], d[0], d[4], 1.0f + d[5], d[1], 0.0f, 0.0f, 1.0f}};
        const float detdW = dW.m[0] * dW.m[4] - dW.m[1] * dW.m[3];
        bool fin = true;
        for (int q = 0; q < 6; ++q) fin = fin && finite_bits(d[q]);
        if (!fin || !(std::fabs(detdW) > 1e-6f)) return reject(iter + 1);
        const M3 Wn = mul3(W, inverse3(dW));
        for (int q = 0; q < 9; ++q)
            if (!finite_bits(Wn.m[q])) return reject(iter + 1);
        W = Wn;

        float nrm = 0.0f;
        for (int q = 0; q < 6; ++q) nrm += d[q] * d[q];
        if (std::sqrt(nrm) < conv_tol) {
            float sgx = 0.0f, sgy = 0.0f;
            int surv = 0;
            for (int k = 0; k < n; ++k) {
                if (!valid[k]) continue;
                sgx += s.gxv[k] * s.gxv[k];
                sgy += s.gyv[k] * s.gyv[k];
                surv++;
            }
            if (surv == 0) return reject(iter + 1);
            if (std::min(sgx, sgy) <= 0.0f) return reject(iter + 1);
            if (!finite_bits(final_score)) return reject(iter + 1);
            return result_from(W, 0, final_score, iter + 1, invalid_ref);
        }
    }
    return result_from(W, 1, final_score, kIcgnMaxIter, invalid_ref);
}

// optimization_engine.cpp:467-528.
float Engine::evaluate_znssd(const Subset& s, const Image& def, const float p[6]) {
    const int n = s.n;
    float mean = 0.0f;
    const int nv = sample_w(*this, s, def, warp_matrix(p), buf_simplex, val_simplex, mean);
    if (nv == 0 || nv < (int)((float)n * kMinValidPixelFraction)) return 2.0f;
    mean /= (float)nv;
    float z = 0.0f;
    if (nv == n) {
        const float ssq = sum_sq_diff(buf_simplex.data(), n, mean);
        float dstd = std::sqrt(ssq / nv);
        if (dstd < 1e-5f) dstd = 1.0f;
        z = znssd_sum(buf_simplex.data(), s.nref.data(), n, mean, 1.0f / dstd);
    } else {
        float ssq = 0.0f;
        for (int i = 0; i < n; ++i)
            if (val_simplex[i]) {
                const float d = buf_simplex[i] - mean;
                ssq += d * d;
            }
        float dstd = std::sqrt(ssq / nv);
        if (dstd < 1e-5f) dstd = 1.0f;
        for (int i = 0; i < n; ++i)
            if (val_simplex[i]) {
                const float nd = (buf_simplex[i] - mean) / dstd;
                const float d = s.nref[i] - nd;
                z += d * d;
            }
    }
    const float score = z / (float)nv;
    return finite_bits(score) ? score : 2.0f;
}

// optimization_engine.cpp:531-651.
Result Engine::solve_simplex(const Subset& s, const Image& def, const Result& start, bool trans) {
    const int D = trans ? 2 : 6;
    const int np = D + 1;
    float p[7][6] = {};
    float y[7] = {};
    const float scale[6] = {2.0f, 2.0f, 0.01f, 0.01f, 0.01f, 0.01f};
    auto eval = [&](const float* pt) {
        float q[6] = {pt[0], pt[1], 0.0f, 0.0f, 0.0f, 0.0f};
        if (!trans)
            for (int j = 2; j < 6; ++j) q[j] = pt[j];
        return evaluate_znssd(s, def, q);
    };
    p[0][0] = start.u;
    p[0][1] = start.v;
    if (!trans) {
        p[0][2] = start.ux;
        p[0][3] = start.uy;
        p[0][4] = start.vx;
        p[0][5] = start.vy;
    }
    y[0] = eval(p[0]);
    for (int i = 1; i < np; ++i) {
        for (int j = 0; j < D; ++j) p[i][j] = p[0][j];
        p[i][i - 1] += scale[i - 1];
        y[i] = eval(p[i]);
    }
    const float alpha = 1.0f, gamma = 2.0f, rho = 0.5f, sigma = 0.5f;
    for (int it = 0; it < 100; ++it) {
        int idx[7] = {0, 1, 2, 3, 4, 5, 6};
        for (int a = 1; a < np; ++a) {
            const int v = idx[a];
            int b = a;
            while (b > 0 && y[v] < y[idx[b - 1]]) {
                idx[b] = idx[b - 1];
                --b;
            }
            idx[b] = v;
        }
        if (std::fabs(y[idx[0]] - y[idx[np - 1]]) < 1e-5f) break;
        float pb[6] = {0, 0, 0, 0, 0, 0};
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) pb[j] += p[idx[i]][j];
        for (int j = 0; j < D; ++j) pb[j] /= D;
        float pr[6] = {0, 0, 0, 0, 0, 0};
        for (int j = 0; j < D; ++j) pr[j] = pb[j] + alpha * (pb[j] - p[idx[np - 1]][j]);
        const float yr = eval(pr);
        if (y[idx[0]] <= yr && yr < y[idx[np - 2]]) {
            for (int j = 0; j < D; ++j) p[idx[np - 1]][j] = pr[j];
            y[idx[np - 1]] = yr;
        } else if (yr < y[idx[0]]) {
            float pe[6] = {0, 0, 0, 0, 0, 0};
            for (int j = 0; j < D; ++j) pe[j] = pb[j] + gamma * (pr[j] - pb[j]);
            const float ye = eval(pe);
            const float* src = (ye < yr) ? pe : pr;
            for (int j = 0; j < D; ++j) p[idx[np - 1]][j] = src[j];
            y[idx[np - 1]] = (ye < yr) ? ye : yr;
        } else {
            float pc[6] = {0, 0, 0, 0, 0, 0};
            const bool outside = yr < y[idx[np - 1]];
            const float* base = outside ? pr : p[idx[np - 1]];
            for (int j = 0; j < D; ++j) pc[j] = pb[j] + rho * (base[j] - pb[j]);
            const float yc = eval(pc);
            if (yc < std::min(yr, y[idx[np - 1]])) {
                for (int j = 0; j < D; ++j) p[idx[np - 1]][j] = pc[j];
                y[idx[np - 1]] = yc;
            } else {
                for (int i = 1; i < np; ++i) {
                    for (int j = 0; j < D; ++j)
                        p[idx[i]][j] = p[idx[0]][j] + sigma * (p[idx[i]][j] - p[idx[0]][j]);
                    y[idx[i]] = eval(p[idx[i]]);
                }
            }
        }
    }
    int best = 0;
    for (int k = 1; k < np; ++k)
        if (y[k] < y[best]) best = k;
    Result r;
    r.u = p[best][0];
    r.v = p[best][1];
    if (!trans) {
        r.ux = p[best][2];
        r.uy = p[best][3];
        r.vx = p[best][4];
        r.vy = p[best][5];
    }
    r.status = (y[best] > kSimplexFailScore) ? -2 : 0;
    r.corr = y[best];
    return r;
}

// ---------------------------------------------------------------------------
// FP64 IC-GN of the same model (no reference counterpart: the error yardstick).
// ---------------------------------------------------------------------------
namespace {

struct Img64 {
    int w, h;
    const std::vector<float>* I;
    double at(int x, int y) const { return (double)(*I)[(size_t)y * w + x]; }
};

double bilinear64(const Img64& m, double x, double y) {
    if (!(x >= 0.0 && x < m.w - 1.5 && y >= 0.0 && y < m.h - 1.5)) return 0.0;
    const int xi = (int)x, yi = (int)y;
    const double dx = x - xi, dy = y - yi;
    const double a = m.at(xi, yi) * (1 - dx) + m.at(xi + 1, yi) * dx;
    const double b = m.at(xi, yi + 1) * (1 - dx) + m.at(xi + 1, yi + 1) * dx;
    return a * (1 - dy) + b * dy;
}

void keys4_64(double s, double* w) {
    const double s2 = s * s, s3 = s2 * s;
    w[0] = -0.5 * s3 + s2 - 0.5 * s;
    w[1] = 1.5 * s3 - 2.5 * s2 + 1.0;
    w[2] = -1.5 * s3 + 2.0 * s2 + 0.5 * s;
    w[3] = 0.5 * s3 - 0.5 * s2;
}
double k0(double s) { return (4.0 / 3.0) * s * s * s - (7.0 / 3.0) * s * s + 1.0; }
double k1(double s) { return -(7.0 / 12.0) * s * s * s + 3.0 * s * s - (59.0 / 12.0) * s + 2.5; }
double k2(double s) { return (1.0 / 12.0) * s * s * s - (2.0 / 3.0) * s * s + (7.0 / 4.0) * s - 1.5; }

double interp64(const Img64& m, double x, double y, bool use6) {
    if (use6) {
        if (!(x > 2.5 && x < m.w - 3.5 && y > 2.5 && y < m.h - 3.5)) return bilinear64(m, x, y);
        const int xi = (int)x, yi = (int)y;
        const double dx = x - xi, dy = y - yi;
        const double wx[6] = {k2(dx + 2), k1(dx + 1), k0(dx), k0(1 - dx), k1(2 - dx), k2(3 - dx)};
        const double wy[6] = {k2(dy + 2), k1(dy + 1), k0(dy), k0(1 - dy), k1(2 - dy), k2(3 - dy)};
        double v = 0;
        for (int j = 0; j < 6; ++j) {
            double r = 0;
            for (int i = 0; i < 6; ++i) r += m.at(xi + i - 2, yi + j - 2) * wx[i];
            v += r * wy[j];
        }
        return v;
    }
    if (!(x >= 1.0 && x < m.w - 2.0 && y >= 1.0 && y < m.h - 2.0)) return bilinear64(m, x, y);
    const int xi = (int)x, yi = (int)y;
    double wx[4], wy[4];
    keys4_64(x - xi, wx);
    keys4_64(y - yi, wy);
    double v = 0;
    for (int j = 0; j < 4; ++j) {
        double r = 0;
        for (int i = 0; i < 4; ++i) r += m.at(xi + i - 1, yi + j - 1) * wx[i];
        v += r * wy[j];
    }
    return v;
}

}  // namespace

Result64 icgn_fp64(const Image& ref, const Image& def, int cx, int cy, int dim, bool use6,
                   const double g[6], double tol, int max_iter) {
    Result64 out;
    const int half = dim / 2, n = dim * dim;
    if (cx - half < 2 || cy - half < 2 || cx + half >= ref.w - 2 || cy + half >= ref.h - 2)
        return out;
    std::vector<double> r(n), gxs(n), gys(n), nref(n);
    std::vector<unsigned char> rv(n);
    const Img64 R{ref.w, ref.h, &ref.I};
    const Img64 D{def.w, def.h, &def.I};
    double sum = 0;
    int nvr = 0;
    for (int i = 0; i < n; ++i) {
        const int x = cx + i % dim - half, y = cy + i / dim - half;
        r[i] = R.at(x, y);
        rv[i] = ref.I[(size_t)y * ref.w + x] >= kGhostWallThreshold;
        // gradients of the same float plane, in double
        gxs[i] = (-R.at(x + 2, y) + 8.0 * R.at(x + 1, y) - 8.0 * R.at(x - 1, y) + R.at(x - 2, y)) / 12.0;
        gys[i] = (-R.at(x, y + 2) + 8.0 * R.at(x, y + 1) - 8.0 * R.at(x, y - 1) + R.at(x, y - 2)) / 12.0;
        if (rv[i]) {
            sum += r[i];
            nvr++;
        }
    }
    if (nvr < 3) return out;
    const double mean = sum / nvr;
    double ss = 0;
    for (int i = 0; i < n; ++i)
        if (rv[i]) ss += (r[i] - mean) * (r[i] - mean);
    double sd = std::sqrt(ss / nvr);
    if (sd < 1e-12) sd = 1.0;
    std::vector<double> sdi((size_t)6 * n);
    for (int i = 0; i < n; ++i) {
        const double ox = i % dim - half, oy = i / dim - half;
        nref[i] = (r[i] - mean) / sd;
        const double a = gxs[i] / sd, b = gys[i] / sd;
        const double s6[6] = {a, b, a * ox, a * oy, b * ox, b * oy};
        for (int q = 0; q < 6; ++q) sdi[(size_t)q * n + i] = s6[q];
    }
    double p[6];
    std::memcpy(p, g, sizeof p);
    std::vector<double> vals(n);
    std::vector<unsigned char> val(n);
    for (int it = 0; it < max_iter; ++it) {
        const double W00 = 1 + p[2], W01 = p[3], W02 = p[0], W10 = p[4], W11 = 1 + p[5], W12 = p[1];
        int nv = 0;
        double dsum = 0;
        for (int i = 0; i < n; ++i) {
            const double ox = i % dim - half, oy = i / dim - half;
            const double X = cx + W00 * ox + W01 * oy + W02;
            const double Y = cy + W10 * ox + W11 * oy + W12;
            const double rx = X + 0.5, ry = Y + 0.5;
            val[i] = 0;
            if (!(rx >= 4.0 && rx < def.w - 4 && ry >= 4.0 && ry < def.h - 4)) continue;
            vals[i] = interp64(D, X, Y, use6);
            if (!rv[i]) continue;
            val[i] = 1;
            dsum += vals[i];
            nv++;
        }
        if (nv < (int)(n * 0.9)) return out;
        const double dm = dsum / nv;
        double dss = 0;
        for (int i = 0; i < n; ++i)
            if (val[i]) dss += (vals[i] - dm) * (vals[i] - dm);
        double dsd = std::sqrt(dss / nv);
        if (dsd < 1e-12) dsd = 1.0;
        double H[36] = {}, gr[6] = {}, err = 0;
        for (int i = 0; i < n; ++i) {
            if (!val[i]) continue;
            const double d = nref[i] - (vals[i] - dm) / dsd;
            err += d * d;
            for (int a = 0; a < 6; ++a) {
                const double sa = sdi[(size_t)a * n + i];
                gr[a] += sa * d;
                for (int b = 0; b < 6; ++b) H[a * 6 + b] += sa * sdi[(size_t)b * n + i];
            }
        }
        double Hi[36];
        invert6d(H, Hi);
        double dp[6], nrm = 0;
        for (int a = 0; a < 6; ++a) {
            double acc = 0;
            for (int b = 0; b < 6; ++b) acc += Hi[a * 6 + b] * gr[b];
            dp[a] = -acc;
            nrm += dp[a] * dp[a];
        }
        // W <- W * dW^-1 (inverse-compositional update) in double
        const double a = 1 + dp[2], b = dp[3], c = dp[0], d = dp[4], e = 1 + dp[5], f = dp[1];
        const double det = a * e - b * d;
        if (!(std::fabs(det) > 1e-12)) return out;
        const double ia = e / det, ib = -b / det, id = -d / det, ie = a / det;
        const double ic = -(ia * c + ib * f), iff = -(id * c + ie * f);
        const double N00 = W00 * ia + W01 * id, N01 = W00 * ib + W01 * ie, N02 = W00 * ic + W01 * iff + W02;
        const double N10 = W10 * ia + W11 * id, N11 = W10 * ib + W11 * ie, N12 = W10 * ic + W11 * iff + W12;
        p[0] = N02;
        p[1] = N12;
        p[2] = N00 - 1;
        p[3] = N01;
        p[4] = N10;
        p[5] = N11 - 1;
        out.iters = it + 1;
        out.corr = err / nv;
        if (std::sqrt(nrm) < tol) {
            out.status = 0;
            break;
        }
    }
    std::memcpy(out.p, p, sizeof p);
    return out;
}

}  // namespace orc
