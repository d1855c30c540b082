// The subset solver of sdk_solve_nodes for callers that cannot allocate device
// memory (code under HIP graph capture, serving loops): every input and output
// is the caller's device memory, and the solver's own buffers -- the reference
// planes, the per-node pool entries, tasks and results, the hand-off list of
// the fast path and its counter -- live in one fixed device array per GPU.
//
// The array is a file-scope __device__ symbol. Host code never passes the
// symbol itself to a kernel (on the host its name is only a stand-in); the
// first call on each device looks up that device's address of it with
// hipGetSymbolAddress, with the device current, and every later call on the
// device uses the address it keeps.
//
// Calls that overlap on one GPU share the array and race: callers serialise
// them (one stream, or one call at a time).
#include <mutex>

#include "noalloc.hpp"
#include "solver.hpp"

namespace sdic {

namespace {

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kPlaneBytes = align256(sizeof(float) * kNoAllocMaxPixels);
constexpr size_t kOffI = 0;
constexpr size_t kOffGx = kOffI + kPlaneBytes;
constexpr size_t kOffGy = kOffGx + kPlaneBytes;
constexpr size_t kOffPool = kOffGy + kPlaneBytes;
constexpr size_t kOffTasks = kOffPool + align256(sizeof(PoolEntry) * kNoAllocMaxNodes);
constexpr size_t kOffOut = kOffTasks + align256(sizeof(Task) * kNoAllocMaxNodes);
constexpr size_t kOffBail = kOffOut + align256(sizeof(SolveOut) * kNoAllocMaxNodes);
constexpr size_t kOffCount = kOffBail + align256(sizeof(int32_t) * kNoAllocMaxNodes);
constexpr size_t kScratchBytes = kOffCount + 256;
static_assert(kScratchBytes == kNoAllocScratchBytes, "the scratch size the header states");

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// Task list from the caller's node arrays (the host loop of sdk_solve_nodes).
__global__ void k_na_tasks(const int32_t* __restrict__ cxy, const float* __restrict__ guess,
                           const int32_t* __restrict__ modes, int n, Task* __restrict__ tasks) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Task t = {};
    t.cx = cxy[2 * i];
    t.cy = cxy[2 * i + 1];
    t.pool = i;
    t.mode = modes[i];
#pragma unroll
    for (int k = 0; k < 6; ++k) t.g[k] = guess[6 * i + k];
    t.node = i;
    tasks[i] = t;
}

// Results into the caller's outputs, in sdk_solve_nodes' layout.
__global__ void k_na_unpack(const SolveOut* __restrict__ res, int n, int32_t* __restrict__ iout,
                            float* __restrict__ fout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const SolveOut o = res[i];
    iout[5 * i] = o.init;
    iout[5 * i + 1] = o.init ? o.status : 1;
    iout[5 * i + 2] = o.iters;
    iout[5 * i + 3] = o.trigger_iters;
    iout[5 * i + 4] = o.simplex;
#pragma unroll
    for (int k = 0; k < 6; ++k) fout[7 * i + k] = o.init ? o.p[k] : 0.0f;
    fout[7 * i + 6] = o.init ? o.corr : 0.0f;
}

// Results cleared before the solve, so a call never reads the last call's.
__global__ void k_na_clear(SolveOut* __restrict__ res, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) res[i] = SolveOut{};
}

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// This device's address of `scratch`, looked up once per device.
unsigned char* scratch_here() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return nullptr;
        g_ptr[dev] = p;
    }
    return (unsigned char*)g_ptr[dev];
}

}  // namespace

int solve_nodes_noalloc(const float* ref, const void* def, int def_is_f32, int w, int h, int dim, int use6, int lm,
                        float tol, int n, const int32_t* cxy, const float* guess, const int32_t* modes,
                        int32_t* iout, float* fout, hipStream_t st) {
    if (!ref || !def || w <= 0 || h <= 0 || n < 0 || ppl_for(dim) < 0) return kNoAllocErrArgs;
    if (n > 0 && (!cxy || !guess || !modes || !iout || !fout)) return kNoAllocErrArgs;
    if ((size_t)w * (size_t)h > kNoAllocMaxPixels || n > kNoAllocMaxNodes) return kNoAllocErrCapacity;
    if (n == 0) return 0;
    unsigned char* s = scratch_here();
    if (!s) return kNoAllocErrArgs;
    float* I = (float*)(s + kOffI);
    float* gx = (float*)(s + kOffGx);
    float* gy = (float*)(s + kOffGy);
    PoolEntry* pool = (PoolEntry*)(s + kOffPool);
    Task* tasks = (Task*)(s + kOffTasks);
    SolveOut* res = (SolveOut*)(s + kOffOut);
    int* bail = (int*)(s + kOffBail);
    int* nbail = (int*)(s + kOffCount);
    const dim3 blk(256), grd((unsigned)((n + 255) / 256));
    hipError_t e = hipSuccess;
    auto chk = [&e](hipError_t r) {
        if (e == hipSuccess) e = r;
    };
    hipLaunchKernelGGL(k_na_tasks, grd, blk, 0, st, cxy, guess, modes, n, tasks);
    chk(hipGetLastError());
    hipLaunchKernelGGL(k_na_clear, grd, blk, 0, st, res, n);
    chk(hipGetLastError());
    chk(launch_prep_ref_f32(ref, w, h, I, gx, gy, st));
    chk(launch_pool_list(I, gx, gy, w, h, cxy, n, dim, lm != 0, pool, st));
    SolveArgs a = {};
    a.tasks = tasks;
    a.out = res;
    a.nitems = n;
    a.bail = bail;
    a.nbail = nbail;   // zeroed by launch_solve's fill kernel
    a.pool = pool;
    a.refI = I;
    a.refGx = gx;
    a.refGy = gy;
    a.def = def;
    a.w = w;
    a.h = h;
    a.dim = dim;
    a.lm = lm;
    a.tol = tol;
    chk(launch_solve(a, use6 != 0, def_is_f32 != 0, st));
    hipLaunchKernelGGL(k_na_unpack, grd, blk, 0, st, res, n, iout, fout);
    chk(hipGetLastError());
    return e == hipSuccess ? 0 : kNoAllocErrArgs;
}

}  // namespace sdic
