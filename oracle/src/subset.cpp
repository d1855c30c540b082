// CPU ORACLE — test infrastructure only (see oracle.hpp).
// Subset statistics, 6x6 Hessian pool entry and SD images, restating
// src/math/subset_precomputer.cpp and src/math/hessian.hpp of the reference.
#include <limits>

#include "oracle.hpp"

namespace orc {

namespace {

// Gauss-Jordan with partial (row) pivoting. Eigen's PartialPivLU::inverse()
// is the reference's routine; the rounding differs in the last bits, which the
// golden fixtures' tolerances cover. A zero pivot gives a non-finite inverse,
// as Eigen's does, so the callers' finiteness checks decide.
template <typename T>
bool gauss_jordan6(const T* A, T* out) {
    T m[6][12];
    for (int r = 0; r < 6; ++r) {
        for (int c = 0; c < 6; ++c) {
            m[r][c] = A[r * 6 + c];
            m[r][6 + c] = (r == c) ? T(1) : T(0);
        }
    }
    bool ok = true;
    for (int k = 0; k < 6; ++k) {
        int piv = k;
        T best = std::fabs(m[k][k]);
        for (int r = k + 1; r < 6; ++r) {
            if (std::fabs(m[r][k]) > best) {
                best = std::fabs(m[r][k]);
                piv = r;
            }
        }
        if (piv != k)
            for (int c = 0; c < 12; ++c) std::swap(m[k][c], m[piv][c]);
        const T p = m[k][k];
        if (p == T(0)) ok = false;
        const T ip = T(1) / p;
        for (int c = 0; c < 12; ++c) m[k][c] *= ip;
        for (int r = 0; r < 6; ++r) {
            if (r == k) continue;
            const T f = m[r][k];
            if (f == T(0)) continue;
            for (int c = 0; c < 12; ++c) m[r][c] -= f * m[k][c];
        }
    }
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) out[r * 6 + c] = m[r][6 + c];
    return ok;
}

bool all_finite(const float* m, int n) {
    for (int i = 0; i < n; ++i)
        if (!finite_bits(m[i])) return false;
    return true;
}

float norm1(const float* m) {
    float best = 0.0f;
    for (int c = 0; c < 6; ++c) {
        float s = 0.0f;
        for (int r = 0; r < 6; ++r) s += std::fabs(m[r * 6 + c]);
        if (s > best) best = s;
    }
    return best;
}

bool dim_ok(int dim) { return dim >= 3 && dim % 2 == 1; }

bool on_image(const Image& img, int cx, int cy, int half) {
    return cx - half >= 0 && cx + half < img.w && cy - half >= 0 && cy + half < img.h;
}

}  // namespace

bool invert6(const float* A, float* Ainv) { return gauss_jordan6<float>(A, Ainv); }
bool invert6d(const double* A, double* Ainv) { return gauss_jordan6<double>(A, Ainv); }

// hessian.hpp:40-52 — rcond of D H D, D = diag(1, 1, 1/h x4), from the
// explicit inverse; a NaN fails.
bool well_conditioned(const float* H, int half) {
    const float s = half > 1 ? 1.0f / (float)half : 1.0f;
    const float d[6] = {1.0f, 1.0f, s, s, s, s};
    float Hn[36], Hi[36];
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) Hn[r * 6 + c] = d[r] * H[r * 6 + c] * d[c];
    if (!all_finite(Hn, 36)) return false;
    invert6(Hn, Hi);
    if (!all_finite(Hi, 36)) return false;
    const float a = norm1(Hn), b = norm1(Hi);
    return 1.0f / (a * b) > kHessianMinRcond;
}

// subset_precomputer.cpp:42-149.
PoolEntry compute_hessian_only(const Image& ref, int cx, int cy, int dim) {
    PoolEntry e;
    if (!dim_ok(dim) || !on_image(ref, cx, cy, dim / 2)) return e;
    const int n = dim * dim, half = dim / 2;

    float sum = 0.0f;
    int nvalid = 0;
    for (int oy = -half; oy <= half; ++oy) {
        const float* row = &ref.I[(size_t)(cy + oy) * ref.w + cx - half];
        for (int ox = 0; ox < dim; ++ox) {
            if (row[ox] < kGhostWallThreshold) continue;
            sum += row[ox];
            nvalid++;
        }
    }
    if (nvalid < n * 0.5f) return e;
    const float mean = sum / (float)nvalid;

    float ss = 0.0f;
    for (int oy = -half; oy <= half; ++oy) {
        const float* row = &ref.I[(size_t)(cy + oy) * ref.w + cx - half];
        for (int ox = 0; ox < dim; ++ox) {
            if (row[ox] < kGhostWallThreshold) continue;
            const float d = row[ox] - mean;
            ss += d * d;
        }
    }
    float sd = std::sqrt(ss / (float)nvalid);
    if (sd < 1e-5f) sd = 1.0f;
    const float inv = 1.0f / sd;
    e.mean = mean;
    e.stdv = sd;

    float h[21] = {};
    for (int oy = -half; oy <= half; ++oy) {
        const size_t base = (size_t)(cy + oy) * ref.w + cx - half;
        const float* row = &ref.I[base];
        const float* grx = &ref.gx[base];
        const float* gry = &ref.gy[base];
        const float fy = (float)oy;
        for (int ox = -half; ox <= half; ++ox) {
            if (row[ox + half] < kGhostWallThreshold) continue;
            const float a = grx[ox + half] * inv;
            const float b = gry[ox + half] * inv;
            const float fx = (float)ox;
            const float s[6] = {a, b, a * fx, a * fy, b * fx, b * fy};
            int k = 0;
            for (int r = 0; r < 6; ++r)
                for (int c = r; c < 6; ++c) h[k++] += s[r] * s[c];
        }
    }
    int k = 0;
    for (int r = 0; r < 6; ++r)
        for (int c = r; c < 6; ++c) {
            e.H[r * 6 + c] = h[k];
            e.H[c * 6 + r] = h[k];
            ++k;
        }
    const float* H = e.H;
    const float det2 = H[0] * H[7] - H[6] * H[1];
    const float nrm2 = H[0] * H[0] + H[1] * H[1] + H[6] * H[6] + H[7] * H[7];
    const float cond2 = (std::fabs(det2) > 1e-12f) ? (nrm2 / std::fabs(det2)) : 1.0e13f;
    if (cond2 > 1.0e12f || !well_conditioned(e.H, half)) {
        std::memset(e.Hinv, 0, sizeof e.Hinv);
        e.valid = false;
        return e;
    }
    invert6(e.H, e.Hinv);
    e.valid = true;
    return e;
}

// subset_precomputer.cpp:159-265.
void precompute_subset_fast(Subset& s, const Image& ref, int cx, int cy, int dim,
                            const PoolEntry& pe) {
    s.cx = cx;
    s.cy = cy;
    if (!dim_ok(dim) || !on_image(ref, cx, cy, dim / 2) || !pe.valid) {
        s.init = false;
        return;
    }
    const int n = dim * dim, half = dim / 2;
    if (s.dim != dim) {
        s.dim = dim;
        s.n = n;
        s.ox.resize(n);
        s.oy.resize(n);
        int i = 0;
        for (int y = -half; y <= half; ++y)
            for (int x = -half; x <= half; ++x) {
                s.ox[i] = (float)x;
                s.oy[i] = (float)y;
                ++i;
            }
        s.ref.resize(n);
        s.nref.resize(n);
        s.gxv.resize(n);
        s.gyv.resize(n);
        s.sd.resize((size_t)6 * n);
    }
    int i = 0;
    for (int oy = 0; oy < dim; ++oy) {
        const size_t base = (size_t)(cy - half + oy) * ref.w + (cx - half);
        for (int ox = 0; ox < dim; ++ox, ++i) {
            s.ref[i] = ref.I[base + ox];
            s.gxv[i] = ref.gx[base + ox];
            s.gyv[i] = ref.gy[base + ox];
        }
    }
    s.mean = pe.mean;
    s.stdv = pe.stdv;
    const float inv = 1.0f / pe.stdv;
    i = 0;
    for (int oy = -half; oy <= half; ++oy) {
        const float fy = (float)oy;
        for (int ox = -half; ox <= half; ++ox, ++i) {
            s.nref[i] = (s.ref[i] - pe.mean) * inv;
            const float a = s.gxv[i] * inv, b = s.gyv[i] * inv, fx = (float)ox;
            s.sd[i] = a;
            s.sd[n + i] = b;
            s.sd[2 * n + i] = a * fx;
            s.sd[3 * n + i] = a * fy;
            s.sd[4 * n + i] = b * fx;
            s.sd[5 * n + i] = b * fy;
        }
    }
    std::memcpy(s.H, pe.H, sizeof s.H);
    std::memcpy(s.Hinv, pe.Hinv, sizeof s.Hinv);
    s.init = true;
}

}  // namespace orc
