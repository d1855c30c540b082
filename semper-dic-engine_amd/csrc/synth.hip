// Synthetic speckle pairs on the device: the benchmark workloads' images and
// engine-generated training / test pairs.
//
// The continuous field S(p) is a sum of Gaussian blobs (sigma U[2,4] px, sign
// +-1, 0.5 per px^2 over the image padded by 16 px), evaluated within 4 sigma
// from a cell index. The deformed image samples S at the reference point p that
// maps to the pixel q: p + U(p) = q, solved by fixed-point iteration in double
// for a displacement field U (affine; sin/cos + Gaussian; Gaussian bumps +
// rotation; uniaxial stretch). S has mean 0 and variance
// rho * pi * E[sigma^2] = 0.5 * pi * 28/3 (a Poisson field of signed blobs whose
// squared profile integrates to pi sigma^2), so both images map S linearly with
// +-3 standard deviations onto [0, 200] (peak 200, clipped beyond), get
// independent Gaussian noise from a counter hash, and are rounded to u8. The
// scale is analytic, so the bytes do not depend on any reduction order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "synth.hpp"

namespace sdic {

namespace {

__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__host__ __device__ inline float hash_u01(uint64_t seed, uint64_t i) {
    return (float)(mix64(seed * 0x632be59bd9b4e019ull + i) >> 40) * (1.0f / 16777216.0f);
}

constexpr float kCell = 16.0f;

struct Blobs {
    const float *bx, *by, *bk, *ba;
    const int *start, *idx;
    float x0, y0;
    int cw, ch;
};

__device__ float blob_eval(const Blobs& b, float x, float y) {
    const int qx = (int)floorf((x - b.x0) / kCell), qy = (int)floorf((y - b.y0) / kCell);
    float s = 0.0f;
    for (int cy = qy - 1; cy <= qy + 1; ++cy) {
        if (cy < 0 || cy >= b.ch) continue;
        for (int cx = qx - 1; cx <= qx + 1; ++cx) {
            if (cx < 0 || cx >= b.cw) continue;
            const int c = cy * b.cw + cx;
            for (int k = b.start[c]; k < b.start[c + 1]; ++k) {
                const int i = b.idx[k];
                const float dx = x - b.bx[i], dy = y - b.by[i];
                const float t = (dx * dx + dy * dy) * b.bk[i];
                if (t < 8.0f) s += b.ba[i] * expf(-t);
            }
        }
    }
    return s;
}

__device__ void disp(const FieldSpec& f, double x, double y, double& u, double& v) {
    const double dx = x - f.cx, dy = y - f.cy;
    switch (f.kind) {
        case kFieldAffine:
            u = f.p[0] + f.p[2] * dx + f.p[3] * dy;
            v = f.p[1] + f.p[4] * dx + f.p[5] * dy;
            return;
        case kFieldSinGauss: {
            const double pi = 3.14159265358979323846;
            u = f.p[0] * sin(2.0 * pi * x / f.p[1]) * cos(2.0 * pi * y / f.p[2]);
            v = f.p[3] * exp(-(dx * dx + dy * dy) / (2.0 * f.p[4] * f.p[4]));
            return;
        }
        case kFieldBumps: {
            const double c = cos(f.p[0]), s = sin(f.p[0]);
            u = (c - 1.0) * dx - s * dy;
            v = s * dx + (c - 1.0) * dy;
            const int nb = (int)f.p[1];
            for (int k = 0; k < nb; ++k) {
                const double* q = f.p + 2 + 5 * k;
                const double ex = x - q[0], ey = y - q[1];
                const double g = exp(-(ex * ex + ey * ey) / (2.0 * q[2] * q[2]));
                u += q[3] * g;
                v += q[4] * g;
            }
            return;
        }
        case kFieldUniaxial:
            u = f.p[0] * dx;
            v = -f.p[1] * f.p[0] * dy;
            return;
        default:
            u = 0.0;
            v = 0.0;
    }
}

__global__ void k_field(Blobs b, FieldSpec f, int w, int h, float* __restrict__ out) {
    const int64_t n = (int64_t)w * h;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double qx = (double)(i % w), qy = (double)(i / w);
        double px = qx, py = qy;
        if (f.kind != kFieldIdentity) {
            for (int it = 0; it < 30; ++it) {
                double u, v;
                disp(f, px, py, u, v);
                const double nx = qx - u, ny = qy - v;
                const double ch = fabs(nx - px) + fabs(ny - py);
                px = nx;
                py = ny;
                if (ch < 1e-9) break;
            }
        }
        out[i] = blob_eval(b, (float)px, (float)py);
    }
}

__global__ void k_quantize(const float* __restrict__ s, int64_t n, float mid, float scale, float peak, float sigma,
                           uint64_t noise_seed, uint8_t* __restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float v = fminf(fmaxf(mid + s[i] * scale, 0.0f), peak);
        if (sigma > 0.0f) {
            const float u1 = fmaxf(hash_u01(noise_seed, 2ull * i), 1e-7f);
            const float u2 = hash_u01(noise_seed, 2ull * i + 1);
            v += sigma * sqrtf(-2.0f * logf(u1)) * cosf(6.283185307f * u2);
        }
        v = fminf(fmaxf(v, 0.0f), 255.0f);
        out[i] = (uint8_t)(v + 0.5f);
    }
}

}  // namespace

// Builds the blob field for a w x h image (host), uploads it, renders the
// reference (identity field) and the deformed image, and writes both as u8
// device planes. Returns a HIP error code (0 on success).
int synth_pair(uint64_t seed, uint64_t noise_seed, int w, int h, const FieldSpec& field, float noise_sigma,
               uint8_t* d_ref, uint8_t* d_def, hipStream_t st) {
    const float pad = 16.0f, density = 0.5f;
    const int nb = (int)std::lround((double)density * (w + 2 * pad) * (h + 2 * pad));
    std::vector<float> bx(nb), by(nb), bk(nb), ba(nb);
    for (int i = 0; i < nb; ++i) {
        bx[i] = -pad + hash_u01(seed, 4ull * i) * (w + 2 * pad);
        by[i] = -pad + hash_u01(seed, 4ull * i + 1) * (h + 2 * pad);
        const float s = 2.0f + 2.0f * hash_u01(seed, 4ull * i + 2);
        bk[i] = 1.0f / (2.0f * s * s);
        ba[i] = hash_u01(seed, 4ull * i + 3) < 0.5f ? -1.0f : 1.0f;
    }
    const float x0 = -pad, y0 = -pad;
    const int cw = (int)std::ceil((w + 2 * pad) / kCell) + 1, ch = (int)std::ceil((h + 2 * pad) / kCell) + 1;
    std::vector<int> cnt((size_t)cw * ch + 1, 0), cell(nb);
    for (int i = 0; i < nb; ++i) {
        const int qx = std::min(cw - 1, std::max(0, (int)((bx[i] - x0) / kCell)));
        const int qy = std::min(ch - 1, std::max(0, (int)((by[i] - y0) / kCell)));
        cell[i] = qy * cw + qx;
        cnt[cell[i] + 1]++;
    }
    for (size_t c = 1; c < cnt.size(); ++c) cnt[c] += cnt[c - 1];
    std::vector<int> idx(nb), fill(cnt.begin(), cnt.end() - 1);
    for (int i = 0; i < nb; ++i) idx[fill[cell[i]]++] = i;

    float *dbx = nullptr, *dS = nullptr;
    int *dstart = nullptr, *didx = nullptr;
    const int64_t n = (int64_t)w * h;
    hipError_t e = hipMalloc(&dbx, sizeof(float) * 4 * (size_t)nb);
    if (e == hipSuccess) e = hipMalloc(&dstart, sizeof(int) * cnt.size());
    if (e == hipSuccess) e = hipMalloc(&didx, sizeof(int) * (size_t)nb);
    if (e == hipSuccess) e = hipMalloc(&dS, sizeof(float) * 2 * (size_t)n);
    if (e == hipSuccess) e = hipMemcpyAsync(dbx, bx.data(), sizeof(float) * nb, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dbx + nb, by.data(), sizeof(float) * nb, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dbx + 2 * (size_t)nb, bk.data(), sizeof(float) * nb, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dbx + 3 * (size_t)nb, ba.data(), sizeof(float) * nb, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dstart, cnt.data(), sizeof(int) * cnt.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(didx, idx.data(), sizeof(int) * nb, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        Blobs b{dbx, dbx + nb, dbx + 2 * (size_t)nb, dbx + 3 * (size_t)nb, dstart, didx, x0, y0, cw, ch};
        FieldSpec ident{};
        ident.kind = kFieldIdentity;
        const int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
        hipLaunchKernelGGL(k_field, dim3(blocks), dim3(256), 0, st, b, ident, w, h, dS);
        hipLaunchKernelGGL(k_field, dim3(blocks), dim3(256), 0, st, b, field, w, h, dS + n);
        const float sd = std::sqrt(density * 3.14159265f * (28.0f / 3.0f)), peak = 200.0f;
        const float scale = 0.5f * peak / (3.0f * sd);
        hipLaunchKernelGGL(k_quantize, dim3(blocks), dim3(256), 0, st, dS, n, 0.5f * peak, scale, peak, noise_sigma,
                           noise_seed ^ 0x5151ull, d_ref);
        hipLaunchKernelGGL(k_quantize, dim3(blocks), dim3(256), 0, st, dS + n, n, 0.5f * peak, scale, peak, noise_sigma,
                           noise_seed ^ 0xa2a2ull, d_def);
        e = hipGetLastError();
        const hipError_t es = hipStreamSynchronize(st);
        if (e == hipSuccess) e = es;
    }
    (void)hipFree(dbx);
    (void)hipFree(dstart);
    (void)hipFree(didx);
    (void)hipFree(dS);
    return (int)e;
}

}  // namespace sdic
