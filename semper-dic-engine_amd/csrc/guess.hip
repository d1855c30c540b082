// Seeding guess field on the grid (build_mesh_guess_field,
// src/pipeline/full_field_mesh.cpp:32-282), one thread per grid node:
//   1. the first lattice triangle containing the node (pointPolygonTest >= 0);
//   2. with a FULL mesh, the triangle with the largest signed distance within
//      3 steps (extrapolation);
//   3. the global affine model for the rest;
//   4. 5x5 box smoothing over the in-mesh nodes, gradients first, u / v
//      carried back from the box centroid along the smoothed gradients.
// The triangles come from the anchor lattice (host), two per lattice cell
// with four kept corners, one with three.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "solver.hpp"

#pragma clang fp contract(off)

namespace sdic {

namespace {

// pointPolygonTest over a triangle: > 0 inside, 0 on an edge, < 0 outside;
// with measure, the signed distance to the nearest edge. Same as the oracle's.
__device__ double tri_test(const TriDev& t, float x, float y, bool measure) {
    int pos = 0, neg = 0, zero = 0;
    double dmin = 1e300;
    for (int k = 0; k < 3; ++k) {
        const double ax = t.px[k], ay = t.py[k];
        const double bx = t.px[(k + 1) % 3], by = t.py[(k + 1) % 3];
        const double cr = (bx - ax) * ((double)y - ay) - (by - ay) * ((double)x - ax);
        if (cr > 0) pos++;
        else if (cr < 0) neg++;
        else zero++;
        if (measure) {
            const double ex = bx - ax, ey = by - ay;
            const double l2 = ex * ex + ey * ey;
            double tt = l2 > 0 ? ((x - ax) * ex + (y - ay) * ey) / l2 : 0.0;
            tt = fmax(0.0, fmin(1.0, tt));
            const double qx = ax + tt * ex - x, qy = ay + tt * ey - y;
            dmin = fmin(dmin, sqrt(qx * qx + qy * qy));
        }
    }
    const bool inside = (pos == 0 || neg == 0);
    const bool on_edge = inside && zero > 0;
    if (!measure) return on_edge ? 0.0 : (inside ? 1.0 : -1.0);
    if (on_edge) return 0.0;
    return inside ? dmin : -dmin;
}

struct Planes {
    uint8_t* in_mesh;
    float *u, *v, *ux, *uy, *vx, *vy;
};

__device__ void assign(const Planes& p, GridDev& g, int idx, const TriDev& t, float X, float Y, int tag) {
    p.u[idx] = (float)(t.ux * X + t.uy * Y + t.u);
    p.v[idx] = (float)(t.vx * X + t.vy * Y + t.v);
    p.ux[idx] = (float)t.ux;
    p.uy[idx] = (float)t.uy;
    p.vx[idx] = (float)t.vx;
    p.vy[idx] = (float)t.vy;
    p.in_mesh[idx] = 1;
    g.tag[idx] = (uint8_t)tag;
}

__global__ void k_guess(GuessArgs a, GridDev g, Planes p) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.gw * a.gh) return;
    p.u[idx] = a.gU;
    p.v[idx] = a.gV;
    p.ux[idx] = p.uy[idx] = p.vx[idx] = p.vy[idx] = 0.0f;
    p.in_mesh[idx] = 0;
    if (g.solved[idx]) return;
    const int gx = idx % a.gw, gy = idx / a.gw;
    const float X = (float)(a.rect_x + gx * a.step), Y = (float)(a.rect_y + gy * a.step);
    // 1. containing triangle: only the cells whose closed rectangle holds the
    //    node, in cell order (a triangle lies inside its cell)
    if (a.LW >= 2 && a.LH >= 2) {
        for (int lj = 0; lj + 1 < a.LH; ++lj) {
            if (gy < a.lat_y[lj] || gy > a.lat_y[lj + 1]) continue;
            for (int li = 0; li + 1 < a.LW; ++li) {
                if (gx < a.lat_x[li] || gx > a.lat_x[li + 1]) continue;
                const int c = lj * (a.LW - 1) + li;
                for (int k = a.cell_tri[c]; k < a.cell_tri[c + 1]; ++k) {
                    const TriDev& t = a.tris[k];
                    if (X < t.bx0 || X > t.bx1 || Y < t.by0 || Y > t.by1) continue;
                    if (tri_test(t, X, Y, false) >= 0) {
                        assign(p, g, idx, t, X, Y, kMeshInTriangle);
                        return;
                    }
                }
            }
        }
        // 2. extrapolation within 3 steps of a triangle (FULL meshes only)
        if (a.quality == 2) {
            const float lim = -3.0f * a.step;
            float best = lim - 1.0f;
            int bt = -1;
            // a triangle lies in its cell, its box padded by 15 px: only cells
            // within 15 px + 3 steps of the node can hold one that qualifies;
            // visiting them in cell order keeps the triangles' order
            const float reach = 15.0f - lim;
            for (int lj = 0; lj + 1 < a.LH; ++lj) {
                if ((float)(a.rect_y + a.lat_y[lj + 1] * a.step) + reach < Y ||
                    (float)(a.rect_y + a.lat_y[lj] * a.step) - reach > Y)
                    continue;
                for (int li = 0; li + 1 < a.LW; ++li) {
                    if ((float)(a.rect_x + a.lat_x[li + 1] * a.step) + reach < X ||
                        (float)(a.rect_x + a.lat_x[li] * a.step) - reach > X)
                        continue;
                    const int c = lj * (a.LW - 1) + li;
                    for (int k = a.cell_tri[c]; k < a.cell_tri[c + 1]; ++k) {
                        const TriDev& t = a.tris[k];
                        if (X < t.bx0 + lim || X > t.bx1 - lim || Y < t.by0 + lim || Y > t.by1 - lim) continue;
                        const double d = tri_test(t, X, Y, true);
                        if (d >= lim && (float)d > best) {
                            best = (float)d;
                            bt = k;
                        }
                    }
                }
            }
            if (bt >= 0) {
                assign(p, g, idx, a.tris[bt], X, Y, kMeshExtrapolated);
                return;
            }
        }
    }
    // 3. the global model
    if (a.have_affine) {
        const double Xd = a.rect_x + (double)gx * a.step, Yd = a.rect_y + (double)gy * a.step;
        const double dx = Xd - a.acx, dy = Yd - a.acy;
        p.u[idx] = (float)(a.aff[0] + a.aff[1] * dx + a.aff[2] * dy);
        p.v[idx] = (float)(a.aff[3] + a.aff[4] * dx + a.aff[5] * dy);
        p.ux[idx] = (float)a.aff[1];
        p.uy[idx] = (float)a.aff[2];
        p.vx[idx] = (float)a.aff[4];
        p.vy[idx] = (float)a.aff[5];
        p.in_mesh[idx] = 1;
        g.tag[idx] = kMeshHomography;
    }
}

// 4a. box average of the four gradient planes (each from its own copy)
__global__ void k_smooth_grad(int gw, int gh, const uint8_t* __restrict__ in_mesh, const float* __restrict__ src,
                              float* __restrict__ dst) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = gw * gh;
    if (idx >= n) return;
    const int x = idx % gw, y = idx / gw;
    for (int f = 0; f < 4; ++f) {
        const float* s = src + (size_t)f * n;
        float* d = dst + (size_t)f * n;
        if (!in_mesh[idx]) {
            d[idx] = s[idx];
            continue;
        }
        float sum = 0.0f;
        int cnt = 0;
        for (int dy = -2; dy <= 2; ++dy)
            for (int dx = -2; dx <= 2; ++dx) {
                const int ny = y + dy, nx = x + dx;
                if (nx >= 0 && nx < gw && ny >= 0 && ny < gh && in_mesh[ny * gw + nx]) {
                    sum += s[ny * gw + nx];
                    cnt++;
                }
            }
        d[idx] = cnt ? sum / cnt : s[idx];
    }
}

// 4b. u / v: box average carried back from the centroid along the smoothed gradients
__global__ void k_smooth_uv(int gw, int gh, float stepf, const uint8_t* __restrict__ in_mesh,
                            const float* __restrict__ src, float* __restrict__ dst, const float* __restrict__ grads) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = gw * gh;
    if (idx >= n) return;
    const int x = idx % gw, y = idx / gw;
    for (int f = 0; f < 2; ++f) {
        const float* s = src + (size_t)f * n;
        float* d = dst + (size_t)f * n;
        const float* gxp = grads + (size_t)(2 * f) * n;       // ux | vx
        const float* gyp = grads + (size_t)(2 * f + 1) * n;   // uy | vy
        if (!in_mesh[idx]) {
            d[idx] = s[idx];
            continue;
        }
        float sum = 0.0f;
        int cnt = 0, sdx = 0, sdy = 0;
        for (int dy = -2; dy <= 2; ++dy)
            for (int dx = -2; dx <= 2; ++dx) {
                const int ny = y + dy, nx = x + dx;
                if (nx >= 0 && nx < gw && ny >= 0 && ny < gh && in_mesh[ny * gw + nx]) {
                    sum += s[ny * gw + nx];
                    cnt++;
                    sdx += dx;
                    sdy += dy;
                }
            }
        if (cnt == 0) {
            d[idx] = s[idx];
            continue;
        }
        float val = sum / cnt;
        if (sdx != 0 || sdy != 0) val -= (gxp[idx] * sdx + gyp[idx] * sdy) * stepf / cnt;
        d[idx] = val;
    }
}

}  // namespace

// Planes gu..gvy must be consecutive in the order u, v, ux, uy, vx, vy with
// stride gw*gh (one allocation); tmp holds 6 planes.
hipError_t launch_guess(const GuessArgs& a, GridDev g, uint8_t* in_mesh, float* gu, float* gv, float* gux,
                        float* guy, float* gvx, float* gvy, float* tmp, hipStream_t st) {
    const int n = a.gw * a.gh;
    if (n <= 0) return hipSuccess;
    const size_t N = (size_t)n;
    if (gv != gu + N || gux != gu + 2 * N || guy != gu + 3 * N || gvx != gu + 4 * N || gvy != gu + 5 * N)
        return hipErrorInvalidValue;
    Planes p{in_mesh, gu, gv, gux, guy, gvx, gvy};
    const dim3 b(256), gr((n + 255) / 256);
    hipLaunchKernelGGL(k_guess, gr, b, 0, st, a, g, p);
    // gradients (planes 2..5) -> tmp[2..5]; then u, v with them -> tmp[0..1]
    hipLaunchKernelGGL(k_smooth_grad, gr, b, 0, st, a.gw, a.gh, in_mesh, gu + 2 * N, tmp + 2 * N);
    // tmp gradient order is ux, uy, vx, vy: u uses (ux, uy) = planes 0,1 of grads, v uses (vx, vy) = 2,3
    hipLaunchKernelGGL(k_smooth_uv, gr, b, 0, st, a.gw, a.gh, (float)a.step, in_mesh, gu, tmp, tmp + 2 * N);
    return hipMemcpyAsync(gu, tmp, 6 * N * sizeof(float), hipMemcpyDeviceToDevice, st);
}

}  // namespace sdic
