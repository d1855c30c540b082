// CPU ORACLE — test infrastructure only (see oracle.hpp).
//
// Two synthetic-image generators:
//  * the reference test suite's own speckle field (tests/framework/synthetic.h:35-187,
//    synthetic_cv.h:21-36): a seeded mt19937 sum of 400/600 Gaussian blobs on a
//    128 base, clamped to [5, 250], sampled at the analytic inverse warp. The
//    reference's golden fixtures were captured on these exact images, so the
//    draws (mt19937 + uniform_real_distribution<float> + std::exp, libstdc++)
//    are reproduced to regenerate them;
//  * the speckle the benchmark workloads are specified on (seeded blobs,
//    sigma U[2,4] px, 0.5 blobs/px^2, min-max scaled to a peak of 200, 2%
//    Gaussian noise), from a counter-based hash so it is independent of any
//    library RNG. The GPU engine has its own generator of the same field; the
//    tests hand the same bytes to both sides, so bit equality of the two
//    generators is not required.
#include <algorithm>
#include <random>

#include "oracle.hpp"
#include "synth.hpp"

namespace orc {

// ----- reference test-suite field (synthetic.h) ------------------------------

RefSpeckle::RefSpeckle(unsigned seed, int width, int height, int blobs) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> ux(0.0f, (float)width);
    std::uniform_real_distribution<float> uy(0.0f, (float)height);
    std::uniform_real_distribution<float> us(1.5f, 3.5f);
    std::uniform_real_distribution<float> ua(-90.0f, 90.0f);
    for (int i = 0; i < blobs; ++i) {
        cx.push_back(ux(rng));
        cy.push_back(uy(rng));
        const float s = us(rng);
        inv2s2.push_back(1.0f / (2.0f * s * s));
        amp.push_back(ua(rng));
    }
}

float RefSpeckle::sample(float x, float y) const {
    float v = 128.0f;
    for (size_t i = 0; i < cx.size(); ++i) {
        const float dx = x - cx[i], dy = y - cy[i];
        const float r2 = dx * dx + dy * dy;
        if (r2 < 40.0f) v += amp[i] * std::exp(-r2 * inv2s2[i]);
    }
    if (v < 5.0f) v = 5.0f;
    if (v > 250.0f) v = 250.0f;
    return v;
}

void RefSpeckle::render(int w, int h, const float* aff, float* out) const {
    // aff: u, v, ux, uy, vx, vy, cx, cy (or null for the reference image)
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            float rx = (float)x, ry = (float)y;
            if (aff) {
                const float bx = (float)x - aff[6] - aff[0];
                const float by = (float)y - aff[7] - aff[1];
                const float a11 = 1.0f + aff[2], a12 = aff[3];
                const float a21 = aff[4], a22 = 1.0f + aff[5];
                const float det = a11 * a22 - a12 * a21;
                const float px = (a22 * bx - a12 * by) / det;
                const float py = (-a21 * bx + a11 * by) / det;
                rx = aff[6] + px;
                ry = aff[7] + py;
            }
            out[(size_t)y * w + x] = sample(rx, ry);
        }
    }
}

void gray8(const float* in, int w, int h, float noise_sigma, unsigned noise_seed, uint8_t* out) {
    std::mt19937 rng(noise_seed);
    std::normal_distribution<float> gauss(0.0f, noise_sigma);
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        float v = in[i];
        if (noise_sigma > 0.0f) v += gauss(rng);
        if (v < 0.f) v = 0.f;
        if (v > 255.f) v = 255.f;
        out[i] = (uint8_t)(v + 0.5f);
    }
}

// ----- benchmark speckle (counter hash) --------------------------------------

static inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

float hash_u01(uint64_t seed, uint64_t i) {
    return (float)(mix64(seed * 0x632be59bd9b4e019ull + i) >> 40) * (1.0f / 16777216.0f);
}

BlobField::BlobField(uint64_t seed, int width, int height, float density, float pad) {
    w = width;
    h = height;
    const int nb = (int)std::lround((double)density * (width + 2 * pad) * (height + 2 * pad));
    bx.resize(nb);
    by.resize(nb);
    bk.resize(nb);
    ba.resize(nb);
    for (int i = 0; i < nb; ++i) {
        bx[i] = -pad + hash_u01(seed, 4ull * i) * (width + 2 * pad);
        by[i] = -pad + hash_u01(seed, 4ull * i + 1) * (height + 2 * pad);
        const float s = 2.0f + 2.0f * hash_u01(seed, 4ull * i + 2);
        bk[i] = 1.0f / (2.0f * s * s);
        ba[i] = hash_u01(seed, 4ull * i + 3) < 0.5f ? -1.0f : 1.0f;
    }
    // bucket by cells of kCell px
    x0 = -pad;
    y0 = -pad;
    cw = (int)std::ceil((width + 2 * pad) / kCell) + 1;
    ch = (int)std::ceil((height + 2 * pad) / kCell) + 1;
    std::vector<int> cnt((size_t)cw * ch + 1, 0);
    std::vector<int> cell(nb);
    for (int i = 0; i < nb; ++i) {
        const int qx = std::min(cw - 1, std::max(0, (int)((bx[i] - x0) / kCell)));
        const int qy = std::min(ch - 1, std::max(0, (int)((by[i] - y0) / kCell)));
        cell[i] = qy * cw + qx;
        cnt[cell[i] + 1]++;
    }
    for (size_t c = 1; c < cnt.size(); ++c) cnt[c] += cnt[c - 1];
    start = cnt;
    idx.resize(nb);
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int i = 0; i < nb; ++i) idx[fill[cell[i]]++] = i;
}

// Sum of the blobs within 4 sigma (sigma <= 4: 16 px), visiting the 3x3 cells
// around the point, in blob index order within each cell.
float BlobField::eval(float x, float y) const {
    const int qx = (int)std::floor((x - x0) / kCell), qy = (int)std::floor((y - y0) / kCell);
    float s = 0.0f;
    for (int cy_ = qy - 1; cy_ <= qy + 1; ++cy_) {
        if (cy_ < 0 || cy_ >= ch) continue;
        for (int cx_ = qx - 1; cx_ <= qx + 1; ++cx_) {
            if (cx_ < 0 || cx_ >= cw) continue;
            const int c = cy_ * cw + cx_;
            for (int k = start[c]; k < start[c + 1]; ++k) {
                const int i = idx[k];
                const float dx = x - bx[i], dy = y - by[i];
                const float r2 = dx * dx + dy * dy;
                const float t = r2 * bk[i];
                if (t < 8.0f) s += ba[i] * std::exp(-t);
            }
        }
    }
    return s;
}

}  // namespace orc
