// CPU ORACLE — test infrastructure only (see oracle.hpp). Synthetic images.
#pragma once

#include <cstdint>
#include <vector>

namespace orc {

// The reference test suite's continuous speckle (tests/framework/synthetic.h).
struct RefSpeckle {
    std::vector<float> cx, cy, inv2s2, amp;
    RefSpeckle(unsigned seed, int width, int height, int blobs);
    float sample(float x, float y) const;
    // aff = {u, v, ux, uy, vx, vy, cx, cy}: deformed image by the analytic
    // inverse warp; null: the reference image.
    void render(int w, int h, const float* aff, float* out) const;
};

// synthetic_cv.h:21-36 — clamp in float, optional mt19937 noise, v + 0.5 truncated.
void gray8(const float* in, int w, int h, float noise_sigma, unsigned noise_seed, uint8_t* out);

// Benchmark speckle: blob i at hash-uniform position over the padded image,
// sigma U[2,4], sign +-1. eval() is the unnormalised sum S(x, y).
float hash_u01(uint64_t seed, uint64_t i);
struct BlobField {
    static constexpr float kCell = 16.0f;
    int w = 0, h = 0;
    std::vector<float> bx, by, bk, ba;
    float x0 = 0, y0 = 0;
    int cw = 0, ch = 0;
    std::vector<int> start, idx;
    BlobField(uint64_t seed, int width, int height, float density, float pad);
    float eval(float x, float y) const;
};

}  // namespace orc
