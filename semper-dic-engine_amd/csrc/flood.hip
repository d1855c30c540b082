// The layered Path B's frontier on the device (see flood.hpp).
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "flood.hpp"
#include "wave.hpp"

#pragma clang fp contract(off)

namespace sdic {

namespace {

// Appends the lanes with `take` to a list through one atomic per wave (the
// same-address atomics of one per lane would serialize at the L2); returns the
// lane's slot, or -1. Every lane of the wave must call it.
__device__ __forceinline__ int wave_append(int* counter, bool take) {
    const uint64_t m = __ballot(take);
    if (m == 0) return -1;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, leader);
    return take ? base + __popcll(m & ((1ull << lane) - 1ull)) : -1;
}

// Adds a per-lane count into a counter through one atomic per wave.
__device__ __forceinline__ void wave_add(int* counter, int v) {
    const int s = wave_sum_i(v);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(counter, s);
}

// A search key read at the L2, where the other waves' atomics left it.
__device__ __forceinline__ unsigned long long key_at(const unsigned long long* k, int i) {
    return __hip_atomic_load(k + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

constexpr int kDx[4] = {1, -1, 0, 0}, kDy[4] = {0, 0, 1, -1};   // right, left, down, up

__global__ void k_flood_init(FloodDev f, const uint8_t* __restrict__ accepted, const uint8_t* __restrict__ solved,
                             int reach) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kFcCount) f.counters[i] = 0;
    if (i >= f.gw * f.gh) return;
    f.ok[i] = accepted[i];
    f.fin[i] = solved[i];
    f.tried[i] = (int8_t)(reach + 1);
    f.committed[i] = 0;
    f.stamp[i] = -1;
    f.cand[i] = -1;
    f.dist[i] = 0;
    f.key[i] = ~0ull;
}

__global__ void k_flood_zero(int* c, int from, int to) {
    const int i = from + (int)threadIdx.x;
    if (i < to) c[i] = 0;
}

// The accepted nodes next to open ones, each with its sort key (corr, node):
// corr is >= 0 for an accepted node, so its bits order as the floats do (a
// negative zero is taken as zero).
__global__ void k_flood_sources(FloodDev f) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int NN = f.gw * f.gh;
    bool take = false;
    if (j < NN && f.ok[j]) {
        const int x = j % f.gw, y = j / f.gw;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int nx = x + kDx[k], ny = y + kDy[k];
            if (nx >= 0 && nx < f.gw && ny >= 0 && ny < f.gh && !f.fin[ny * f.gw + nx]) take = true;
        }
    }
    const int pos = wave_append(f.counters + kFcSources, take);
    if (take) {
        const float c = f.w[(size_t)6 * NN + j] + 0.0f;
        f.skeys[pos] = ((unsigned long long)__float_as_uint(c) << 32) | (unsigned)j;
    }
}

__global__ void k_flood_unpack(FloodDev f, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) f.sources[t] = (int)(f.skeys_sorted[t] & 0xffffffffull);
}

// The search and the layer's tasks, one workgroup.
__global__ __launch_bounds__(1024) void k_flood_search(FloodDev f, int ns, int reach, int tag, int rect_x,
                                                       int rect_y, int step, int learned) {
    __shared__ int len_next, reached_n;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int NN = f.gw * f.gh;
    for (int t = tid; t < ns; t += nt) f.key[f.sources[t]] = (unsigned long long)t << 32;
    if (tid == 0) reached_n = 0;
    __syncthreads();
    const int* cur = f.sources;
    int len = ns;
    for (int d = 0; d < reach && len > 0; ++d) {
        const int shift = 30 - 2 * d, ltag = tag * 32 + d;
        int* nxt = (d & 1) ? f.level_b : f.level_a;
        if (tid == 0) len_next = 0;
        __syncthreads();
        for (int idx = tid; idx < len; idx += nt) {
            const int p = cur[idx];
            const unsigned long long kp = key_at(f.key, p);
            const int x = p % f.gw, y = p / f.gw;
            for (int k = 0; k < 4; ++k) {
                const int nx = x + kDx[k], ny = y + kDy[k];
                if (nx < 0 || nx >= f.gw || ny < 0 || ny >= f.gh) continue;
                const int m = ny * f.gw + nx;
                if (f.fin[m] || f.ok[m] || f.stamp[m] == tag) continue;
                atomicMin(f.key + m, kp | ((unsigned long long)k << shift));
                if (atomicExch(f.cand + m, ltag) != ltag) nxt[atomicAdd(&len_next, 1)] = m;
            }
        }
        __syncthreads();
        const int nb = len_next, r0 = reached_n;
        for (int idx = tid; idx < nb; idx += nt) {
            const int m = nxt[idx];
            f.stamp[m] = tag;
            f.dist[m] = d + 1;
            f.reached[r0 + idx] = m;
        }
        __syncthreads();
        if (tid == 0) reached_n = r0 + nb;
        cur = nxt;
        len = nb;
        __syncthreads();
    }
    const int R = reached_n;
    // one task per reached node whose claim comes closer than its last
    // failed one: the seed's warp carried over the offset
    for (int base = 0; base < R; base += nt) {
        const int idx = base + tid;
        bool take = false;
        int i = -1;
        if (idx < R) {
            i = f.reached[idx];
            take = f.dist[i] < f.tried[i];
        }
        const int pos = wave_append(f.counters + kFcTasks, take);
        if (take) {
            const unsigned long long ki = key_at(f.key, i);
            const int s0 = f.sources[(int)(ki >> 32)], x = i % f.gw, y = i / f.gw;
            const float ddx = (float)((x - s0 % f.gw) * step), ddy = (float)((y - s0 / f.gw) * step);
            const float* w = f.w;
            Task tk;
            tk.cx = rect_x + x * step;
            tk.cy = rect_y + y * step;
            tk.pool = i;
            tk.node = i;
            tk.pad = 0;
            tk.mode = (f.dist[i] == 1 && !learned) ? NO_SEARCH : NO_SIMPLEX;
            tk.g[0] = w[s0] + w[(size_t)2 * NN + s0] * ddx + w[(size_t)3 * NN + s0] * ddy;
            tk.g[1] = w[(size_t)NN + s0] + w[(size_t)4 * NN + s0] * ddx + w[(size_t)5 * NN + s0] * ddy;
            for (int c = 0; c < 4; ++c) tk.g[2 + c] = w[(size_t)(2 + c) * NN + s0];
            f.tasks[pos] = tk;
            f.tdist[pos] = f.dist[i];
        }
    }
    __syncthreads();
    // the keys back to unset for the next layer
    for (int idx = tid; idx < R; idx += nt) f.key[f.reached[idx]] = ~0ull;
    for (int t = tid; t < ns; t += nt) f.key[f.sources[t]] = ~0ull;
    if (tid == 0) f.counters[kFcReached] = R;
}

__global__ void k_flood_update(FloodDev f, const SolveOut* __restrict__ res, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int NN = f.gw * f.gh;
    int points = 0, iters = 0, calls = 0, saved = 0, dead = 0;
    if (t < n) {
        const SolveOut r = res[t];
        const int i = f.tasks[t].node;
        const bool acc = r.init && r.status == 0 && r.corr <= kCorrAccept;
        if (r.init) {
            f.commit_res[i] = r;
            f.committed[i] = 1;
            iters = r.iters;
            if (r.simplex) {
                calls = 1;
                saved = acc ? 1 : 0;
                dead = acc ? 0 : 1;
            }
        }
        if (acc) {
            points = 1;
            f.ok[i] = 1;
            f.fin[i] = 1;
            for (int c = 0; c < 6; ++c) f.w[(size_t)c * NN + i] = r.p[c];
            f.w[(size_t)6 * NN + i] = r.corr;
        } else if (f.tdist[t] == 1 || !r.init) {
            f.fin[i] = 1;
        } else {
            f.tried[i] = 2;
        }
    }
    wave_add(f.counters + kFcPoints, points);
    wave_add(f.counters + kFcIters, iters);
    wave_add(f.counters + kFcCalls, calls);
    wave_add(f.counters + kFcSaved, saved);
    wave_add(f.counters + kFcDead, dead);
}

__global__ void k_flood_commit_list(FloodDev f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool take = i < f.gw * f.gh && f.committed[i];
    const int pos = wave_append(f.counters + kFcCommitted, take);
    if (take) {
        f.commit_nodes[pos] = i;
        f.commit_list[pos] = f.commit_res[i];
    }
}

}  // namespace

hipError_t launch_flood_init(const FloodDev& f, const uint8_t* accepted, const uint8_t* solved, const float* planes,
                             int reach, hipStream_t st) {
    const int NN = f.gw * f.gh;
    hipError_t e = hipMemcpyAsync(f.w, planes, sizeof(float) * 7 * (size_t)NN, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
    const int n = NN > kFcCount ? NN : kFcCount;
    hipLaunchKernelGGL(k_flood_init, dim3((n + 255) / 256), dim3(256), 0, st, f, accepted, solved, reach);
    return hipGetLastError();
}

hipError_t launch_flood_sources(const FloodDev& f, hipStream_t st) {
    const int NN = f.gw * f.gh;
    hipLaunchKernelGGL(k_flood_zero, dim3(1), dim3(64), 0, st, f.counters, (int)kFcSources, (int)kFcTasks + 1);
    hipLaunchKernelGGL(k_flood_sources, dim3((NN + 255) / 256), dim3(256), 0, st, f);
    return hipGetLastError();
}

size_t flood_sort_temp_bytes(int n) {
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, (const unsigned long long*)nullptr,
                                            (unsigned long long*)nullptr, n, 0, 64);
    return bytes;
}

hipError_t launch_flood_sort(const FloodDev& f, int nsources, void* temp, size_t temp_bytes, hipStream_t st) {
    if (nsources <= 0) return hipSuccess;
    hipError_t e = hipcub::DeviceRadixSort::SortKeys(temp, temp_bytes, f.skeys, f.skeys_sorted, nsources, 0, 64, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_flood_unpack, dim3((nsources + 255) / 256), dim3(256), 0, st, f, nsources);
    return hipGetLastError();
}

hipError_t launch_flood_search(const FloodDev& f, int nsources, int reach, int tag, int rect_x, int rect_y, int step,
                               bool learned, hipStream_t st) {
    if (reach < 1 || reach > kFloodDeviceMaxReach || tag < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_flood_search, dim3(1), dim3(1024), 0, st, f, nsources, reach, tag, rect_x, rect_y, step,
                       learned ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_flood_update(const FloodDev& f, const SolveOut* res, int ntasks, hipStream_t st) {
    if (ntasks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_flood_update, dim3((ntasks + 255) / 256), dim3(256), 0, st, f, res, ntasks);
    return hipGetLastError();
}

hipError_t launch_flood_commit_list(const FloodDev& f, hipStream_t st) {
    const int NN = f.gw * f.gh;
    hipLaunchKernelGGL(k_flood_zero, dim3(1), dim3(64), 0, st, f.counters, (int)kFcCommitted, (int)kFcCommitted + 1);
    hipLaunchKernelGGL(k_flood_commit_list, dim3((NN + 255) / 256), dim3(256), 0, st, f);
    return hipGetLastError();
}

}  // namespace sdic
