// CPU ORACLE — test infrastructure only (see oracle.hpp).
// Image planes, 4th-order gradients and interpolation, restating
// src/math/image_processor.cpp of the reference.
#include "oracle.hpp"

namespace orc {

// image_processor.cpp:9-24 — u8 to float, exact.
void image_from_u8(Image& img, int w, int h, const uint8_t* px) {
    img.w = w;
    img.h = h;
    img.I.resize((size_t)w * h);
    for (size_t i = 0; i < img.I.size(); ++i) img.I[i] = (float)px[i];
    image_gradients(img);
}

void image_from_f32(Image& img, int w, int h, const float* px) {
    img.w = w;
    img.h = h;
    img.I.assign(px, px + (size_t)w * h);
    image_gradients(img);
}

// image_processor.cpp:54-63 — (-I[+2] + 8 I[+1] - 8 I[-1] + I[-2]) / 12, a 2 px
// border left at zero.
void image_gradients(Image& img) {
    const int w = img.w, h = img.h;
    img.gx.assign((size_t)w * h, 0.0f);
    img.gy.assign((size_t)w * h, 0.0f);
    const float* I = img.I.data();
    for (int y = 2; y < h - 2; ++y) {
        for (int x = 2; x < w - 2; ++x) {
            const int i = y * w + x;
            img.gx[i] = (-I[i + 2] + 8.0f * I[i + 1] - 8.0f * I[i - 1] + I[i - 2]) / 12.0f;
            img.gy[i] = (-I[i + 2 * w] + 8.0f * I[i + w] - 8.0f * I[i - w] + I[i - 2 * w]) / 12.0f;
        }
    }
}

// image_processor.cpp:198-212 — 2x2, zero outside [0, dim - 1.5).
float interp_bilinear(const Image& img, float x, float y) {
    if (!(x >= 0.0f && x < img.w - 1.5f && y >= 0.0f && y < img.h - 1.5f)) return 0.0f;
    const int xi = (int)x, yi = (int)y;
    const float dx = x - xi, dy = y - yi;
    const float* r0 = &img.I[(size_t)yi * img.w + xi];
    const float* r1 = r0 + img.w;
    const float a = r0[0] * (1.0f - dx) + r0[1] * dx;
    const float b = r1[0] * (1.0f - dx) + r1[1] * dx;
    return a * (1.0f - dy) + b * dy;
}

// Keys cubic convolution, a = -0.5 (image_processor.cpp:188-195).
static inline void keys4_weights(float s, float* w) {
    const float s2 = s * s;
    const float s3 = s2 * s;
    w[0] = -0.5f * s3 + s2 - 0.5f * s;
    w[1] = 1.5f * s3 - 2.5f * s2 + 1.0f;
    w[2] = -1.5f * s3 + 2.0f * s2 + 0.5f * s;
    w[3] = 0.5f * s3 - 0.5f * s2;
}

// image_processor.cpp:214-242 — demotes to bilinear outside [1, dim - 2).
float interp_bicubic(const Image& img, float x, float y) {
    if (!(x >= 1.0f && x < img.w - 2.0f && y >= 1.0f && y < img.h - 2.0f))
        return interp_bilinear(img, x, y);
    const int xi = (int)x, yi = (int)y;
    float wx[4], wy[4];
    keys4_weights(x - xi, wx);
    keys4_weights(y - yi, wy);
    float val = 0.0f;
    for (int j = 0; j < 4; ++j) {
        const float* r = &img.I[(size_t)(yi + j - 1) * img.w + xi];
        float rv = 0.0f;
        rv += r[-1] * wx[0];
        rv += r[0] * wx[1];
        rv += r[1] * wx[2];
        rv += r[2] * wx[3];
        val += rv * wy[j];
    }
    return val;
}

// Keys 4th-order piecewise cubics on |s| in [0,1], (1,2], (2,3]
// (image_processor.cpp:71-82).
static inline float k6_f0(float s) { return (4.0f / 3.0f) * s * s * s - (7.0f / 3.0f) * s * s + 1.0f; }
static inline float k6_f1(float s) {
    return -(7.0f / 12.0f) * s * s * s + 3.0f * s * s - (59.0f / 12.0f) * s + 2.5f;
}
static inline float k6_f2(float s) {
    return (1.0f / 12.0f) * s * s * s - (2.0f / 3.0f) * s * s + (7.0f / 4.0f) * s - 1.5f;
}

// image_processor.cpp:84-123 — demotes to bilinear outside (2.5, dim - 3.5).
float interp_keys6(const Image& img, float x, float y) {
    if (!(x > 2.5f && x < img.w - 3.5f && y > 2.5f && y < img.h - 3.5f))
        return interp_bilinear(img, x, y);
    const int xi = (int)x, yi = (int)y;
    const float dx = x - xi, dy = y - yi;
    const float wx[6] = {k6_f2(dx + 2.0f), k6_f1(dx + 1.0f), k6_f0(dx),
                         k6_f0(1.0f - dx), k6_f1(2.0f - dx), k6_f2(3.0f - dx)};
    const float wy[6] = {k6_f2(dy + 2.0f), k6_f1(dy + 1.0f), k6_f0(dy),
                         k6_f0(1.0f - dy), k6_f1(2.0f - dy), k6_f2(3.0f - dy)};
    float val = 0.0f;
    for (int j = 0; j < 6; ++j) {
        const float* r = &img.I[(size_t)(yi + j - 2) * img.w + xi];
        float rv = 0.0f;
        rv += r[-2] * wx[0];
        rv += r[-1] * wx[1];
        rv += r[0] * wx[2];
        rv += r[1] * wx[3];
        rv += r[2] * wx[4];
        rv += r[3] * wx[5];
        val += rv * wy[j];
    }
    return val;
}

}  // namespace orc
