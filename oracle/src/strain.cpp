// CPU ORACLE — test infrastructure only (see oracle.hpp).
// VSG strain and the normalized median test, restating
// src/strain/strain_calculator.cpp and src/util/median.hpp of the reference.
#include <algorithm>

#include "oracle.hpp"

namespace orc {

// src/util/median.hpp:21-30 — NaNs dropped, even count: mean of the middle pair.
float median_inplace(float* v, size_t n) {
    n = (size_t)(std::partition(v, v + n, [](float x) { return x == x; }) - v);
    if (n == 0) return 0.0f;
    const size_t mid = n / 2;
    std::nth_element(v, v + mid, v + n);
    const float hi = v[mid];
    if (n % 2 == 1) return hi;
    const float lo = *std::max_element(v, v + mid);
    return 0.5f * (lo + hi);
}

// strain_calculator.cpp:12-25.
int vsg_window_node_count(int step, int window_px) {
    if (step <= 0) return 0;
    const float radius = window_px / 2.0f;
    const int grid_rad = (int)std::ceil(radius / (float)step);
    const double r2 = (double)(radius * radius) + kVsgRadiusTiny;
    int count = 0;
    for (int dy = -grid_rad; dy <= grid_rad; ++dy)
        for (int dx = -grid_rad; dx <= grid_rad; ++dx) {
            const double px = (double)(dx * step), py = (double)(dy * step);
            if (px * px + py * py <= r2) count++;
        }
    return count;
}

// strain_calculator.cpp:27-158 — FP64 plane fit over the circular window,
// >= 90% fill and >= 3 nodes, 1-norm rcond >= 1e-12, Green-Lagrange strain.
StrainField compute_vsg_strain(const DispField& d, int window_px) {
    StrainField s;
    const size_t want = d.w > 0 && d.h > 0 ? (size_t)d.w * d.h : 0;
    s.exx.assign(want, kStrainUninit);
    s.eyy.assign(want, kStrainUninit);
    s.exy.assign(want, kStrainUninit);
    if (d.step <= 0 || window_px <= 0 || d.u.size() < want || d.v.size() < want) return s;
    const float radius = window_px / 2.0f;
    const float radius_sq = radius * radius;
    const int grid_rad = (int)std::ceil(radius / d.step);
    const int expected = vsg_window_node_count(d.step, window_px);
    const double r2 = (double)radius_sq + kVsgRadiusTiny;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < d.h; ++y) {
        for (int x = 0; x < d.w; ++x) {
            const int idx = y * d.w + x;
            if (!d.valid[idx]) continue;
            double A[9] = {}, bu[3] = {}, bv[3] = {};
            int np = 0;
            for (int dy = -grid_rad; dy <= grid_rad; ++dy)
                for (int dx = -grid_rad; dx <= grid_rad; ++dx) {
                    const int nx = x + dx, ny = y + dy;
                    if (nx < 0 || nx >= d.w || ny < 0 || ny >= d.h) continue;
                    const int nidx = ny * d.w + nx;
                    if (!d.valid[nidx]) continue;
                    const double px = (double)(dx * d.step), py = (double)(dy * d.step);
                    if (px * px + py * py > r2) continue;
                    const double a[3] = {1.0, px, py};
                    for (int r = 0; r < 3; ++r) {
                        for (int c = 0; c < 3; ++c) A[r * 3 + c] += a[r] * a[c];
                        bu[r] += a[r] * (double)d.u[nidx];
                        bv[r] += a[r] * (double)d.v[nidx];
                    }
                    np++;
                }
            const double fill = (double)np / (double)expected;
            if (!(fill >= 0.90 && np >= 3)) continue;
            double anorm = 0.0;
            for (int c = 0; c < 3; ++c)
                anorm = std::max(anorm, std::fabs(A[c]) + std::fabs(A[3 + c]) + std::fabs(A[6 + c]));
            // cofactor inverse with Eigen's default |det| > 1e-12 invertibility test
            auto M = [&](int r, int c) { return A[r * 3 + c]; };
            auto cof = [&](int i, int j) {
                const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
                return M(i1, j1) * M(i2, j2) - M(i1, j2) * M(i2, j1);
            };
            const double det = cof(0, 0) * M(0, 0) + cof(1, 0) * M(1, 0) + cof(2, 0) * M(2, 0);
            if (!(std::fabs(det) > 1e-12)) continue;
            double Ai[9];
            const double inv = 1.0 / det;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) Ai[i * 3 + j] = cof(j, i) * inv;
            double inorm = 0.0;
            for (int c = 0; c < 3; ++c)
                inorm = std::max(inorm, std::fabs(Ai[c]) + std::fabs(Ai[3 + c]) + std::fabs(Ai[6 + c]));
            const double rcond = (anorm * inorm > 0.0) ? 1.0 / (anorm * inorm) : 0.0;
            if (rcond < 1e-12) continue;
            double cu[3], cv[3];
            for (int r = 0; r < 3; ++r) {
                cu[r] = Ai[r * 3] * bu[0] + Ai[r * 3 + 1] * bu[1] + Ai[r * 3 + 2] * bu[2];
                cv[r] = Ai[r * 3] * bv[0] + Ai[r * 3 + 1] * bv[1] + Ai[r * 3 + 2] * bv[2];
            }
            const double ux = cu[1], uy = cu[2], vx = cv[1], vy = cv[2];
            s.exx[idx] = (float)(0.5 * (2.0 * ux + ux * ux + vx * vx));
            s.eyy[idx] = (float)(0.5 * (2.0 * vy + uy * uy + vy * vy));
            s.exy[idx] = (float)(0.5 * (uy + vx + ux * uy + vx * vy));
        }
    }
    return s;
}

namespace {

float normalized_residual(float centre, std::vector<float>& nb, std::vector<float>& scratch) {
    const float med = median_inplace(nb.data(), nb.size());
    scratch.resize(nb.size());
    for (size_t i = 0; i < nb.size(); ++i) scratch[i] = std::fabs(nb[i] - med);
    const float spread = median_inplace(scratch.data(), scratch.size());
    return std::fabs(centre - med) / (spread + kOutlierEpsPx);
}

}  // namespace

// strain_calculator.cpp:174-224 — 5x5, <= 5 passes, each pass reading the
// previous pass's flags.
int reject_displacement_outliers(DispField& d) {
    const int w = d.w, h = d.h, r = kOutlierRadius;
    std::vector<unsigned char> ok = d.valid;
    int rejected = 0;
    for (int pass = 0; pass < kOutlierMaxPasses; ++pass) {
        std::vector<unsigned char> fail((size_t)w * h, 0);
        int nfail = 0;
#pragma omp parallel for schedule(static) reduction(+ : nfail)
        for (int y = 0; y < h; ++y) {
            std::vector<float> nu, nv, scratch;
            nu.reserve(24);
            nv.reserve(24);
            for (int x = 0; x < w; ++x) {
                const int idx = y * w + x;
                if (!ok[idx]) continue;
                nu.clear();
                nv.clear();
                for (int dy = -r; dy <= r; ++dy) {
                    const int ny = y + dy;
                    if (ny < 0 || ny >= h) continue;
                    for (int dx = -r; dx <= r; ++dx) {
                        const int nx = x + dx;
                        if (nx < 0 || nx >= w || (dx == 0 && dy == 0)) continue;
                        const int nidx = ny * w + nx;
                        if (!ok[nidx]) continue;
                        nu.push_back(d.u[nidx]);
                        nv.push_back(d.v[nidx]);
                    }
                }
                if ((int)nu.size() < kOutlierMinNeighbours) continue;
                if (normalized_residual(d.u[idx], nu, scratch) > kOutlierThreshold ||
                    normalized_residual(d.v[idx], nv, scratch) > kOutlierThreshold) {
                    fail[idx] = 1;
                    nfail++;
                }
            }
        }
        if (nfail == 0) break;
        for (size_t i = 0; i < fail.size(); ++i)
            if (fail[i]) ok[i] = 0;
        rejected += nfail;
    }
    for (size_t i = 0; i < ok.size(); ++i)
        if (!ok[i]) d.valid[i] = 0;
    return rejected;
}

}  // namespace orc
