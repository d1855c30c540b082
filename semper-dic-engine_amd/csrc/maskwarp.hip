// Forward warp of the frame-0 material mask for the incremental session
// (warp_material_mask, src/pipeline/incremental_math.cpp:328-447).
//
//   k_mask_warp  one thread per frame-0 pixel of the window: each material
//                pixel is sampled per_px x per_px times, every sample X lands
//                at round(X + U(X)) on the continuous field of the grid
//                (bilinear inside, the plane of the nearest grid point's
//                gradient outside), and the landing pixel is set to 255.
//                Every writer stores the same value, so the result does not
//                depend on the order.
//   k_morph3     3x3 dilation or erosion; pixels outside the image never
//                change the result (the close of a binary mask).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "maskwarp.hpp"

namespace sdic {

namespace {

struct WarpGrid {
    int gw, gh, rx, ry, step;
    const float* U;
    const float* V;
};

// The field and its derivatives at image point (x, y) (ContinuousField::eval,
// incremental_math.cpp:283-322).
__device__ void eval_field(const WarpGrid& g, double x, double y, double& u, double& v) {
    const double gxr = (x - g.rx) / g.step, gyr = (y - g.ry) / g.step;
    const double gx = fmin(fmax(gxr, 0.0), (double)(g.gw - 1));
    const double gy = fmin(fmax(gyr, 0.0), (double)(g.gh - 1));
    const double ox = (gxr - gx) * g.step, oy = (gyr - gy) * g.step;
    int i0 = (int)floor(gx), j0 = (int)floor(gy);
    i0 = max(0, min(i0, max(g.gw - 2, 0)));
    j0 = max(0, min(j0, max(g.gh - 2, 0)));
    const int i1 = min(i0 + 1, g.gw - 1), j1 = min(j0 + 1, g.gh - 1);
    const double tx = (i1 == i0) ? 0.0 : gx - i0;
    const double ty = (j1 == j0) ? 0.0 : gy - j0;
    double val[2], ddx[2], ddy[2];
    const float* F[2] = {g.U, g.V};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double a = F[k][(size_t)j0 * g.gw + i0], b = F[k][(size_t)j0 * g.gw + i1];
        const double c = F[k][(size_t)j1 * g.gw + i0], d = F[k][(size_t)j1 * g.gw + i1];
        val[k] = a * (1 - tx) * (1 - ty) + b * tx * (1 - ty) + c * (1 - tx) * ty + d * tx * ty;
        ddx[k] = (i1 == i0) ? 0.0 : ((b - a) * (1 - ty) + (d - c) * ty) / g.step;
        ddy[k] = (j1 == j0) ? 0.0 : ((c - a) * (1 - tx) + (d - b) * tx) / g.step;
    }
    u = val[0] + ddx[0] * ox + ddy[0] * oy;
    v = val[1] + ddx[1] * ox + ddy[1] * oy;
}

__global__ void k_mask_warp(const uint8_t* __restrict__ mask0, int W, int H, WarpGrid g, int X0, int Y0, int X1,
                            int Y1, int per_px, uint8_t* __restrict__ out) {
    const int ww = X1 - X0 + 1;
    const int64_t n = (int64_t)ww * (Y1 - Y0 + 1);
    const double hs = 1.0 / per_px;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int sx = X0 + (int)(k % ww), sy = Y0 + (int)(k / ww);
        if (mask0[(size_t)sy * W + sx] < 128) continue;
        for (int b = 0; b < per_px; ++b)
            for (int a = 0; a < per_px; ++a) {
                const double X = sx - 0.5 + (a + 0.5) * hs, Y = sy - 0.5 + (b + 0.5) * hs;
                double u, v;
                eval_field(g, X, Y, u, v);
                const long px = lround(X + u), py = lround(Y + v);
                if (px < 0 || px >= W || py < 0 || py >= H) continue;
                out[(size_t)py * W + px] = 255;
            }
    }
}

__global__ void k_morph3(const uint8_t* __restrict__ in, uint8_t* __restrict__ out, int W, int H, int dilate) {
    const int64_t n = (int64_t)W * H;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(k % W), y = (int)(k / W);
        uint8_t r = dilate ? 0 : 255;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int xx = x + dx, yy = y + dy;
                if (xx < 0 || xx >= W || yy < 0 || yy >= H) continue;
                const uint8_t q = in[(size_t)yy * W + xx];
                r = dilate ? (q > r ? q : r) : (q < r ? q : r);
            }
        out[k] = r;
    }
}

int blocks(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (int)(b < 1 ? 1 : b > 65536 ? 65536 : b);
}

template <typename T>
struct Buf {
    T* p = nullptr;
    explicit Buf(size_t n) {
        if (hipMalloc(&p, sizeof(T) * (n ? n : 1)) != hipSuccess) p = nullptr;
    }
    ~Buf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

int mask_warp(const uint8_t* mask0, int W, int H, const MaskGrid& g, const std::vector<float>& U,
              const std::vector<float>& V, int X0, int Y0, int X1, int Y1, int per_px, std::vector<uint8_t>& out,
              hipStream_t st) {
    out.assign((size_t)W * H, 0);
    if (X0 > X1 || Y0 > Y1) return 0;
    const size_t npx = (size_t)W * H, nn = (size_t)g.gw * g.gh;
    Buf<uint8_t> dm(npx), da(npx), db(npx);
    Buf<float> du(nn), dv(nn);
    if (!dm.p || !da.p || !db.p || !du.p || !dv.p) return -3;
    hipError_t e = hipMemcpyAsync(dm.p, mask0, npx, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(du.p, U.data(), sizeof(float) * nn, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dv.p, V.data(), sizeof(float) * nn, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(da.p, 0, npx, st);
    if (e != hipSuccess) return -3;
    const WarpGrid wg{g.gw, g.gh, g.rx, g.ry, g.step, du.p, dv.p};
    const int64_t nwin = (int64_t)(X1 - X0 + 1) * (Y1 - Y0 + 1);
    hipLaunchKernelGGL(k_mask_warp, dim3(blocks(nwin)), dim3(256), 0, st, dm.p, W, H, wg, X0, Y0, X1, Y1, per_px,
                       da.p);
    hipLaunchKernelGGL(k_morph3, dim3(blocks((int64_t)npx)), dim3(256), 0, st, da.p, db.p, W, H, 1);
    hipLaunchKernelGGL(k_morph3, dim3(blocks((int64_t)npx)), dim3(256), 0, st, db.p, da.p, W, H, 0);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out.data(), da.p, npx, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e == hipSuccess ? 0 : -3;
}

}  // namespace sdic
