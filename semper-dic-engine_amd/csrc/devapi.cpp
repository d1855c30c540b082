// Kernel-level C entry points of libsemper_dev.so (declared in
// include/semper_dev.h): the subset solver on float or u8 images, the Hessian
// pool, the image planes, an engine whose frames are already in HBM, device
// buffers and the synthetic speckle generator. The test suite and the
// benchmark use them; applications use the semper_* ABI.
#include <cstring>
#include <memory>
#include <vector>

#include "codec.hpp"
#include "engine.hpp"
#include "handles.hpp"
#include "noalloc.hpp"
#include "semper_c.h"
#include "semper_dev.h"
#include "synth.hpp"

using namespace sdic;

namespace {

template <typename T>
struct Dev {
    T* p = nullptr;
    explicit Dev(size_t n) {
        if (hipMalloc(&p, sizeof(T) * (n ? n : 1)) != hipSuccess) p = nullptr;
    }
    ~Dev() {
        if (p) (void)hipFree(p);
    }
};

struct DevEngine {
    Engine eng;
    Cancel cancel;
};

}  // namespace

extern "C" {

SDK_API int sdk_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

SDK_API void* sdk_malloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) return nullptr;
    return p;
}

SDK_API void sdk_free(void* p) {
    if (p) (void)hipFree(p);
}

SDK_API void* sdk_host_alloc(size_t bytes) {
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) return nullptr;
    return p;
}

SDK_API void sdk_host_free(void* p) {
    if (p) (void)hipHostFree(p);
}

SDK_API int sdk_memcpy(void* dst, const void* src, size_t bytes, int kind) {
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    return (int)hipMemcpy(dst, src, bytes, k);
}

// Subset solver on host images (ref float; def float or u8): the Hessian pool
// of each node centre, then calculate_deformation per node. lm selects the
// full-field engine configuration. Outputs per node: ints {init, status,
// iters, trigger_iters, simplex} and floats {u, v, ux, uy, vx, vy, corr}.
SDK_API int sdk_solve_nodes(const float* ref, const void* def, int def_is_f32, int w, int h, int dim, int use6, int lm,
                            float tol, int n, const int32_t* cxy, const float* guess, const int32_t* modes,
                            int32_t* iout, float* fout) {
    if (!ref || !def || w <= 0 || h <= 0 || n < 0 || ppl_for(dim) < 0) return -3;
    if (n == 0) return 0;
    const size_t npx = (size_t)w * h;
    hipStream_t st;
    if (hipStreamCreate(&st) != hipSuccess) return -3;
    int rc = 0;
    {
        Dev<float> dref(npx), I(npx), gx(npx), gy(npx);
        Dev<uint8_t> ddef(npx * (def_is_f32 ? 4 : 1));
        Dev<int32_t> dcxy(2 * (size_t)n), bail(n), nbail(1);
        Dev<PoolEntry> pool(n);
        Dev<Task> tasks(n);
        Dev<SolveOut> outd(n);
        std::vector<Task> ht(n);
        for (int i = 0; i < n; ++i) {
            Task t = {};
            t.cx = cxy[2 * i];
            t.cy = cxy[2 * i + 1];
            t.pool = i;
            t.mode = modes[i];
            std::memcpy(t.g, guess + 6 * i, sizeof(t.g));
            t.node = i;
            ht[i] = t;
        }
        if (!dref.p || !I.p || !gx.p || !gy.p || !ddef.p || !dcxy.p || !bail.p || !nbail.p || !pool.p || !tasks.p ||
            !outd.p) {
            rc = -3;
        } else {
            hipError_t e = hipMemcpyAsync(dref.p, ref, sizeof(float) * npx, hipMemcpyHostToDevice, st);
            auto chk = [&e](hipError_t r) {
                if (e == hipSuccess) e = r;
            };
            chk(hipMemcpyAsync(ddef.p, def, npx * (def_is_f32 ? 4 : 1), hipMemcpyHostToDevice, st));
            chk(hipMemcpyAsync(dcxy.p, cxy, sizeof(int32_t) * 2 * n, hipMemcpyHostToDevice, st));
            chk(hipMemcpyAsync(tasks.p, ht.data(), sizeof(Task) * n, hipMemcpyHostToDevice, st));
            chk(launch_prep_ref_f32(dref.p, w, h, I.p, gx.p, gy.p, st));
            chk(launch_pool_list(I.p, gx.p, gy.p, w, h, dcxy.p, n, dim, lm != 0, pool.p, st));
            SolveArgs a = {};
            a.tasks = tasks.p;
            a.out = outd.p;
            a.nitems = n;
            a.bail = bail.p;
            a.nbail = nbail.p;
            a.pool = pool.p;
            a.refI = I.p;
            a.refGx = gx.p;
            a.refGy = gy.p;
            a.def = ddef.p;
            a.w = w;
            a.h = h;
            a.dim = dim;
            a.lm = lm;
            a.tol = tol;
            chk(launch_solve(a, use6 != 0, def_is_f32 != 0, st));
            std::vector<SolveOut> ho(n);
            chk(hipMemcpyAsync(ho.data(), outd.p, sizeof(SolveOut) * n, hipMemcpyDeviceToHost, st));
            chk(hipStreamSynchronize(st));
            if (e != hipSuccess) rc = -3;
            for (int i = 0; i < n && rc == 0; ++i) {
                const SolveOut& o = ho[i];
                iout[5 * i] = o.init;
                iout[5 * i + 1] = o.init ? o.status : 1;
                iout[5 * i + 2] = o.iters;
                iout[5 * i + 3] = o.trigger_iters;
                iout[5 * i + 4] = o.simplex;
                const float v[7] = {o.p[0], o.p[1], o.p[2], o.p[3], o.p[4], o.p[5], o.corr};
                for (int k = 0; k < 7; ++k) fout[7 * i + k] = o.init ? v[k] : 0.0f;
            }
        }
    }
    (void)hipStreamDestroy(st);
    return rc;
}

// sdk_solve_nodes without allocation: device pointers, the caller's stream,
// the solver's buffers in the library's fixed device array (csrc/noalloc.hip).
SDK_API int sdk_solve_nodes_noalloc(const float* ref, const void* def, int def_is_f32, int w, int h, int dim, int use6,
                                    int lm, float tol, int n, const int32_t* cxy, const float* guess,
                                    const int32_t* modes, int32_t* iout, float* fout, void* stream) {
    return solve_nodes_noalloc(ref, def, def_is_f32, w, h, dim, use6, lm, tol, n, cxy, guess, modes, iout, fout,
                               (hipStream_t)stream);
}

// Hessian pool entries of a host float image at node centres:
// valid, mean, inv_std and the 36-float step matrix per node.
SDK_API int sdk_pool(const float* ref, int w, int h, int dim, int lm, int n, const int32_t* cxy, int32_t* valid,
                     float* stats2, float* hstep) {
    if (!ref || w <= 0 || h <= 0 || n <= 0) return -3;
    const size_t npx = (size_t)w * h;
    Dev<float> dref(npx), I(npx), gx(npx), gy(npx);
    Dev<int32_t> dcxy(2 * (size_t)n);
    Dev<PoolEntry> pool(n);
    if (!dref.p || !I.p || !gx.p || !gy.p || !dcxy.p || !pool.p) return -3;
    if (hipMemcpy(dref.p, ref, sizeof(float) * npx, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dcxy.p, cxy, sizeof(int32_t) * 2 * n, hipMemcpyHostToDevice) != hipSuccess ||
        launch_prep_ref_f32(dref.p, w, h, I.p, gx.p, gy.p, nullptr) != hipSuccess ||
        launch_pool_list(I.p, gx.p, gy.p, w, h, dcxy.p, n, dim, lm != 0, pool.p, nullptr) != hipSuccess)
        return -3;
    std::vector<PoolEntry> hp(n);
    if (hipMemcpy(hp.data(), pool.p, sizeof(PoolEntry) * n, hipMemcpyDeviceToHost) != hipSuccess) return -3;
    for (int i = 0; i < n; ++i) {
        valid[i] = hp[i].valid;
        stats2[2 * i] = hp[i].mean;
        stats2[2 * i + 1] = hp[i].inv_std;
        std::memcpy(hstep + 36 * i, hp[i].hstep, sizeof(float) * 36);
    }
    return 0;
}

// The reference planes of a u8 image (with an optional ghost-wall mask).
SDK_API int sdk_ref_planes(const uint8_t* img, const uint8_t* mask, int w, int h, float* I, float* gx, float* gy) {
    if (!img || w <= 0 || h <= 0) return -3;
    const size_t npx = (size_t)w * h;
    Dev<uint8_t> di(npx), dm(npx);
    Dev<float> dI(npx), dgx(npx), dgy(npx);
    if (!di.p || !dm.p || !dI.p || !dgx.p || !dgy.p) return -3;
    if (hipMemcpy(di.p, img, npx, hipMemcpyHostToDevice) != hipSuccess) return -3;
    if (mask && hipMemcpy(dm.p, mask, npx, hipMemcpyHostToDevice) != hipSuccess) return -3;
    if (launch_prep_ref_u8(di.p, mask ? dm.p : nullptr, w, h, dI.p, dgx.p, dgy.p, nullptr) != hipSuccess) return -3;
    if (hipMemcpy(I, dI.p, sizeof(float) * npx, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(gx, dgx.p, sizeof(float) * npx, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(gy, dgy.p, sizeof(float) * npx, hipMemcpyDeviceToHost) != hipSuccess)
        return -3;
    return 0;
}

// ---- an engine over frames already in HBM ----------------------------------

SDK_API void* sdk_engine_create(void) {
    try {
        return new DevEngine();
    } catch (...) {
        return nullptr;
    }
}

SDK_API void sdk_engine_destroy(void* h) { delete (DevEngine*)h; }

SDK_API int sdk_engine_set_reference(void* h, const uint8_t* img, int on_device, int w, int hh, const uint8_t* mask) {
    if (!h) return -3;
    return ((DevEngine*)h)->eng.set_reference(img, on_device != 0, w, hh, mask);
}

// params: rect_x, rect_y, rect_w, rect_h, step, subset, strain_window, use6
SDK_API int sdk_engine_run(void* h, const uint8_t* def, int on_device, const int32_t* params, float* out, int cap,
                           float* metrics, int metrics_len, float tol) {
    if (!h || !params) return -3;
    DevEngine* e = (DevEngine*)h;
    Params p;
    p.rect_x = params[0];
    p.rect_y = params[1];
    p.rect_w = params[2];
    p.rect_h = params[3];
    p.step = params[4];
    p.subset = params[5];
    p.strain_window = params[6];
    p.use6 = params[7] != 0;
    e->cancel.enter();
    try {
        return e->eng.run(def, on_device != 0, nullptr, p, out, cap, metrics, metrics_len, e->cancel, nullptr, tol);
    } catch (...) {
        return -3;
    }
}

// Last solve's timings: total, prep, seeding, prepass, mesh, path_a, path_b,
// strain (host ms), the Path A solver launch (device ms), its grid nodes and
// its IC-GN iterations.
// The wave reductions of the solver on nw x 64 host floats: out receives every
// lane's float sum, iout every lane's integer sum of (int)in.
SDK_API int sdk_wave_sum_probe(const float* in, int nw, float* out, int32_t* iout) {
    if (!in || !out || !iout || nw <= 0) return -3;
    const size_t n = (size_t)nw * 64;
    Dev<float> din(n), dout(n);
    Dev<int32_t> di(n);
    if (!din.p || !dout.p || !di.p) return -3;
    if (hipMemcpy(din.p, in, sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess ||
        launch_wave_sum_probe(din.p, nw, dout.p, di.p, nullptr) != hipSuccess ||
        hipMemcpy(out, dout.p, sizeof(float) * n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(iout, di.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost) != hipSuccess)
        return -3;
    return 0;
}

SDK_API int sdk_compose64_probe(const float* planes, const uint8_t* med, int fgw, int fgh, int frx, int fry,
                                int fstep, const float* uf, const float* vf, const double* u64, const double* v64,
                                const uint8_t* alive, int gw, int gh, int rx, int ry, int step, uint8_t* flags,
                                double* ou64, double* ov64, float* ouf, float* ovf, float* ocorr) {
    if (!planes || !med || !uf || !vf || !u64 || !v64 || !alive || !flags || !ou64 || !ov64 || !ouf || !ovf ||
        !ocorr || fgw <= 0 || fgh <= 0 || fstep <= 0 || gw <= 0 || gh <= 0 || step <= 0)
        return -3;
    const size_t NN = (size_t)fgw * fgh, n0 = (size_t)gw * gh;
    Dev<float> dpl(7 * NN), dtf(4 * n0), dcorr(n0);
    Dev<uint8_t> dmed(NN), dfl(2 * n0);
    Dev<double> dt64(4 * n0);
    if (!dpl.p || !dtf.p || !dcorr.p || !dmed.p || !dfl.p || !dt64.p) return -3;
    if (hipMemcpy(dpl.p, planes, sizeof(float) * 7 * NN, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dmed.p, med, NN, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dtf.p, uf, sizeof(float) * n0, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dtf.p + n0, vf, sizeof(float) * n0, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dt64.p, u64, sizeof(double) * n0, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dt64.p + n0, v64, sizeof(double) * n0, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dfl.p, alive, n0, hipMemcpyHostToDevice) != hipSuccess)
        return -3;
    const ComposeField f{dpl.p, dmed.p, fgw, fgh, frx, fry, fstep};
    const ComposeTotals t{dtf.p, dtf.p + n0, dt64.p, dt64.p + n0, dfl.p, gw, gh, rx, ry, step};
    const ComposeOut o{dt64.p + 2 * n0, dt64.p + 3 * n0, dtf.p + 2 * n0, dtf.p + 3 * n0, dcorr.p, dfl.p + n0};
    if (launch_sample_compose64(f, t, o, nullptr) != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(flags, dfl.p + n0, n0, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(ou64, dt64.p + 2 * n0, sizeof(double) * n0, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(ov64, dt64.p + 3 * n0, sizeof(double) * n0, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(ouf, dtf.p + 2 * n0, sizeof(float) * n0, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(ovf, dtf.p + 3 * n0, sizeof(float) * n0, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(ocorr, dcorr.p, sizeof(float) * n0, hipMemcpyDeviceToHost) != hipSuccess)
        return -3;
    return 0;
}

SDK_API int sdk_interp_probe(const uint8_t* img, int w, int h, int form, int n, const float* xs, const float* ys,
                             float* out) {
    if (!img || !xs || !ys || !out || w < 16 || h < 16 || n <= 0 || form < 0 || form > 2 ||
        (size_t)w * h >= kCoefMaxPixels)
        return -3;
    for (int i = 0; i < n; ++i)   // the solver's guard: the stencil lies inside the plane
        if (!(xs[i] >= 4.0f && xs[i] < (float)(w - 5) && ys[i] >= 4.0f && ys[i] < (float)(h - 5))) return -3;
    const size_t npx = (size_t)w * h;
    Dev<uint8_t> d8(npx);
    Dev<float> df(npx), dx(n), dy(n), dout(n);
    Dev<uint4> dc(npx);
    if (!d8.p || !df.p || !dx.p || !dy.p || !dout.p || !dc.p) return -3;
    if (hipMemcpy(d8.p, img, npx, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dx.p, xs, sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dy.p, ys, sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess ||
        launch_u8_to_f32(d8.p, npx, df.p, nullptr) != hipSuccess ||
        launch_row_cubic(d8.p, w, h, dc.p, nullptr) != hipSuccess ||
        launch_interp_probe(df.p, dc.p, d8.p, w, h, form, n, dx.p, dy.p, dout.p, nullptr) != hipSuccess ||
        hipMemcpy(out, dout.p, sizeof(float) * n, hipMemcpyDeviceToHost) != hipSuccess)
        return -3;
    return 0;
}

SDK_API int sdk_engine_set_flag(void* h, int flag, int value) {
    if (!h) return -3;
    switch (flag) {
        case 0:
            ((DevEngine*)h)->eng.speculate_path_b = value != 0;
            return 0;
        case 1:
            ((DevEngine*)h)->eng.out_on_device = value != 0;
            return 0;
        case 2:
            ((DevEngine*)h)->eng.prec.chol64 = value != 0;
            return 0;
        case 3:
            if (value < 0 || value > 2) return -3;
            ((DevEngine*)h)->eng.prec.grad16 = value != 0;
            ((DevEngine*)h)->eng.prec.grad_bf16 = value == 2;
            return 0;
        case 4:
            ((DevEngine*)h)->eng.prec.mfma_pool = value != 0;
            return 0;
        case 5:
            if (value && !((DevEngine*)h)->eng.seeder_ready()) return -3;
            ((DevEngine*)h)->eng.learned_seed = value != 0;
            return 0;
        case 6:
            ((DevEngine*)h)->eng.autotune = value != 0;
            return 0;
        case 7:
            if (value < 0 || value > 6) return -3;
            ((DevEngine*)h)->eng.force_shape = value;
            return 0;
        case 8:
            if (value < 0 || value > Engine::kLayeredReachMax) return -3;
            ((DevEngine*)h)->eng.layered_path_b = value;
            return 0;
        case 9:
            ((DevEngine*)h)->eng.use_graphs = value != 0;
            return 0;
        case 11:
            ((DevEngine*)h)->eng.seeder_mfma = value != 0;
            return 0;
        case 12:
            ((DevEngine*)h)->eng.node_report = value != 0;
            return 0;
        case 13:
            ((DevEngine*)h)->eng.def_f32 = value != 0;
            return 0;
        case 14:
            ((DevEngine*)h)->eng.cnn_rescue = value != 0;
            return 0;
        case 16:
            ((DevEngine*)h)->eng.global_search = value != 0;
            return 0;
        case 17:
            ((DevEngine*)h)->eng.def_coef = value != 0;
            return 0;
        case 18:
            ((DevEngine*)h)->eng.flood_on_device = value != 0;
            return 0;
        case 19:
            ((DevEngine*)h)->eng.cnn_always = value != 0;
            return 0;

        default:
            return -3;
    }
}

static void timings14(const Timings& t, double* out14) {
    const double v[14] = {t.total,  t.prep,   t.seeding,       t.prepass,
                          t.mesh,   t.path_a, t.path_b,        t.strain,
                          t.kernel_path_a,    (double)t.path_a_items, (double)t.path_a_iters,
                          (double)t.path_b_spec, (double)t.path_b_launches, (double)t.path_b_miss};
    std::memcpy(out14, v, sizeof(v));
}

SDK_API int sdk_engine_timings(void* h, double* out14) {
    if (!h || !out14) return -3;
    timings14(((DevEngine*)h)->eng.timings(), out14);
    return 0;
}

// ---- synthetic speckle pairs ------------------------------------------------

// Renders a reference / deformed u8 pair of w x h into device buffers
// (on_device) or host buffers. kind and p[] as FieldSpec (synth.hpp).
SDK_API int sdk_synth_pair(uint64_t seed, int w, int h, int kind, const double* p, int np, double cx, double cy,
                           float noise_sigma, uint8_t* ref, uint8_t* def, int on_device) {
    if (w <= 0 || h <= 0 || np < 0 || np > 48) return -3;
    FieldSpec f = {};
    f.kind = kind;
    f.cx = cx;
    f.cy = cy;
    for (int i = 0; i < np; ++i) f.p[i] = p[i];
    const size_t npx = (size_t)w * h;
    if (on_device) return synth_pair(seed, seed, w, h, f, noise_sigma, ref, def, nullptr);
    Dev<uint8_t> dr(npx), dd(npx);
    if (!dr.p || !dd.p) return -3;
    const int rc = synth_pair(seed, seed, w, h, f, noise_sigma, dr.p, dd.p, nullptr);
    if (rc != 0) return rc;
    if (hipMemcpy(ref, dr.p, npx, hipMemcpyDeviceToHost) != hipSuccess) return -3;
    return hipMemcpy(def, dd.p, npx, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -3;
}

// One deformed frame of a sequence: the speckle of `seed` under the field,
// with its own noise from noise_seed (frames of a sequence do not share a
// noise pattern). Writes w x h bytes to def (host, or device when on_device).
SDK_API int sdk_synth_frame(uint64_t seed, uint64_t noise_seed, int w, int h, int kind, const double* p, int np,
                            double cx, double cy, float noise_sigma, uint8_t* def, int on_device) {
    if (w <= 0 || h <= 0 || np < 0 || np > 48 || !def) return -3;
    FieldSpec f = {};
    f.kind = kind;
    f.cx = cx;
    f.cy = cy;
    for (int i = 0; i < np; ++i) f.p[i] = p[i];
    const size_t npx = (size_t)w * h;
    Dev<uint8_t> dr(npx), dd(on_device ? 1 : npx);
    if (!dr.p || !dd.p) return -3;
    uint8_t* out = on_device ? def : dd.p;
    const int rc = synth_pair(seed, noise_seed, w, h, f, noise_sigma, dr.p, out, nullptr);
    if (rc != 0) return rc;
    if (!on_device && hipMemcpy(def, dd.p, npx, hipMemcpyDeviceToHost) != hipSuccess) return -3;
    return 0;
}

// The boundary's image decode (codec.hpp), for tests of the host side.
SDK_API int sdk_decode(const uint8_t* data, size_t len, int expected_w, int expected_h, int mask, uint8_t* out,
                       size_t cap, int32_t* w, int32_t* h) {
    try {
        std::vector<uint8_t> plane;
        int dw = 0, dh = 0;
        const bool ok = mask ? sdic::codec::decode_mask(data, len, expected_w, expected_h, plane, dw, dh)
                             : sdic::codec::decode_gray(data, len, expected_w, expected_h, plane, dw, dh);
        if (w) *w = ok ? dw : 0;
        if (h) *h = ok ? dh : 0;
        if (!ok) return -3;
        if (out != nullptr) {
            if (cap < plane.size()) return -3;
            std::memcpy(out, plane.data(), plane.size());
        }
        return 0;
    } catch (...) {
        return -3;
    }
}

SDK_API void sdk_image_dimensions(const uint8_t* data, size_t len, int32_t* w, int32_t* h) {
    int dw = 0, dh = 0;
    try {
        sdic::codec::image_dimensions(data, len, dw, dh);
    } catch (...) {
        dw = dh = 0;
    }
    if (w) *w = dw;
    if (h) *h = dh;
}

static int raster_convert(const void* px, int w, int h, int cn, int depth, uint8_t* out, bool mask) {
    using sdic::codec::Depth;
    static const Depth kDepth[6] = {Depth::U8, Depth::U16, Depth::S16, Depth::S32, Depth::F32, Depth::F64};
    static const size_t kBytes[6] = {1, 2, 2, 4, 4, 8};
    if (!px || !out || w <= 0 || h <= 0 || (cn != 1 && cn != 3) || depth < 0 || depth > 5) return -3;
    try {
        sdic::codec::Raster r;
        r.w = w;
        r.h = h;
        r.cn = cn;
        r.depth = kDepth[depth];
        const size_t bytes = (size_t)w * h * cn * kBytes[depth];
        r.px.assign((const uint8_t*)px, (const uint8_t*)px + bytes);
        std::vector<uint8_t> o;
        if (!(mask ? sdic::codec::to_mask8(r, o) : sdic::codec::to_gray8(r, o))) return -3;
        std::memcpy(out, o.data(), o.size());
        return 0;
    } catch (...) {
        return -3;
    }
}

SDK_API int sdk_to_gray8(const void* px, int w, int h, int cn, int depth, uint8_t* out) {
    return raster_convert(px, w, h, cn, depth, out, false);
}

SDK_API int sdk_to_mask8(const void* px, int w, int h, int cn, int depth, uint8_t* out) {
    return raster_convert(px, w, h, cn, depth, out, true);
}

// ---- the learned seeder ------------------------------------------------------

SDK_API int sdk_seeder_param_count(void) { return Seeder::param_count(); }

SDK_API int sdk_seeder_grad(const float* params, const float* x, const float* t, int batch, float* grads,
                            float* loss) {
    if (!params || !x || !t || !grads || !loss) return -3;
    try {
        Seeder s;
        return s.grad(params, x, t, batch, grads, loss);
    } catch (...) {
        return -3;
    }
}

SDK_API int sdk_seeder_forward(const float* params, const float* x, int batch, float* y) {
    if (!params || !x || !y) return -3;
    try {
        Seeder s;
        return s.forward(params, x, batch, y);
    } catch (...) {
        return -3;
    }
}

SDK_API int sdk_seeder_forward_mfma(const float* params, const float* x, int batch, float* y) {
    if (!params || !x || !y) return -3;
    try {
        Seeder s;
        return s.forward_mfma(params, x, batch, y);
    } catch (...) {
        return -3;
    }
}

SDK_API int sdk_seeder_train(uint64_t seed, int steps, int batch, float lr, const char* out_path, float* losses,
                             int nloss) {
    if (!out_path || steps < 1 || batch < 1) return -3;
    try {
        // the reference the pairs are cut from: the engine's own speckle
        // (the workloads' statistics) on 2048 x 2048, noise 4 grey levels
        const int w = 2048, h = 2048;
        Dev<uint8_t> ref((size_t)w * h), tmp((size_t)w * h);
        if (!ref.p || !tmp.p) return -3;
        FieldSpec f = {};
        f.kind = kFieldIdentity;
        if (synth_pair(seed, seed + 1, w, h, f, 4.0f, ref.p, tmp.p, nullptr) != 0) return -3;
        if (hipDeviceSynchronize() != hipSuccess) return -3;
        Seeder s;
        if (s.init(seed) != 0) return -3;
        SeederTrainConfig cfg;
        cfg.steps = steps;
        cfg.batch = batch;
        cfg.lr = lr;
        cfg.seed = seed;
        std::vector<float> hist;
        if (s.train(ref.p, w, h, cfg, &hist) != 0) return -3;
        if (s.save(out_path) != 0) return -3;
        const int k = std::min((int)hist.size(), std::max(nloss, 0));
        if (losses)
            for (int i = 0; i < k; ++i) losses[i] = hist[i];
        return k;
    } catch (...) {
        return -3;
    }
}

SDK_API int sdk_engine_set_region_map(void* h, int rw, int rh, const int32_t* dims) {
    if (!h || (rw * rh > 0 && !dims)) return -3;
    std::vector<int> d;
    if (rw > 0 && rh > 0) d.assign(dims, dims + (size_t)rw * rh);
    return ((DevEngine*)h)->eng.set_region_map(rw, rh, d) == 0 ? 0 : -3;
}

SDK_API int sdk_incremental_set_region_map(semper_incremental* inc, int rw, int rh, const int32_t* dims) {
    if (!inc || (rw * rh > 0 && !dims)) return -3;
    if (inc->busy.exchange(true, std::memory_order_acquire)) return -3;
    std::vector<int> d;
    if (rw > 0 && rh > 0) d.assign(dims, dims + (size_t)rw * rh);
    const int rc = inc->session.set_region_map(rw, rh, d) == 0 ? 0 : -3;
    inc->busy.store(false, std::memory_order_release);
    return rc;
}

SDK_API int sdk_engine_node_report(void* h, uint8_t* tag, uint8_t* path, int n) {
    if (!h || !tag || !path) return -3;
    const Engine& e = ((DevEngine*)h)->eng;
    if (e.report_path.empty() || n < (int)e.report_path.size()) return -3;
    std::memcpy(tag, e.report_tag.data(), e.report_tag.size());
    std::memcpy(path, e.report_path.data(), e.report_path.size());
    return (int)e.report_path.size();
}

SDK_API int sdk_engine_node_planes(void* h, int which, float* out, int n) {
    if (!h || !out || which < 0 || which > 2) return -3;
    const Engine& e = ((DevEngine*)h)->eng;
    const std::vector<float>& v = which == 0 ? e.report_planes : which == 1 ? e.report_planes_a : e.report_guess;
    if (v.empty() || n < (int)v.size()) return -3;
    std::memcpy(out, v.data(), sizeof(float) * v.size());
    return (int)v.size();
}

SDK_API int sdk_engine_global_model(void* h, double* out12) {
    if (!h || !out12) return -3;
    std::memcpy(out12, ((DevEngine*)h)->eng.last_model, sizeof(double) * 12);
    return 0;
}

SDK_API int sdk_engine_graph_stats(void* h, int32_t* stats2) {
    if (!h || !stats2) return -3;
    const auto g = ((DevEngine*)h)->eng.graph_stats();
    stats2[0] = g.captures;
    stats2[1] = g.replays;
    return 0;
}

SDK_API int sdk_incremental_set_flag(semper_incremental* inc, int flag, int value) {
    if (!inc || (flag != 9 && flag != 10)) return -3;
    if (inc->busy.exchange(true, std::memory_order_acquire)) return -3;
    if (flag == 9) inc->session.use_graphs = value != 0;
    else inc->session.compose64 = value != 0;
    inc->busy.store(false, std::memory_order_release);
    return 0;
}

SDK_API int sdk_incremental_timings(semper_incremental* inc, double* t8) {
    if (!inc || !t8) return -3;
    for (int i = 0; i < 8; ++i) t8[i] = inc->session.timing[i];
    return 0;
}

SDK_API int sdk_incremental_engine_timings(semper_incremental* inc, double* out14) {
    if (!inc || !out14 || !inc->session.engine()) return -3;
    timings14(inc->session.engine()->timings(), out14);
    return 0;
}

SDK_API int sdk_incremental_anchor_stats(semper_incremental* inc, int32_t* out2) {
    if (!inc || !out2) return -3;
    out2[0] = inc->session.reanchor_requests;
    out2[1] = inc->session.reanchor_matched;
    return 0;
}

SDK_API int sdk_incremental_graph_stats(semper_incremental* inc, int32_t* stats2) {
    if (!inc || !stats2) return -3;
    const auto g = inc->session.graph_stats();
    stats2[0] = g.captures;
    stats2[1] = g.replays;
    return 0;
}

SDK_API int sdk_engine_tuning(void* h, int32_t* chosen2, double* ms6) {
    if (!h || !chosen2 || !ms6) return -3;
    const auto& r = ((DevEngine*)h)->eng.last_tune;
    if (!r.have) return -3;
    chosen2[0] = r.chosen.wpb;
    chosen2[1] = r.chosen.tile;
    for (int i = 0; i < 6; ++i) ms6[i] = r.ms[i];
    return r.settled ? 0 : 1;
}

SDK_API int sdk_engine_set_subset_map(void* h, int gw, int gh, const int32_t* dims) {
    if (!h) return -3;
    try {
        std::vector<int> d;
        if (dims && gw > 0 && gh > 0) d.assign(dims, dims + (size_t)gw * gh);
        return ((DevEngine*)h)->eng.set_subset_map(gw, gh, d);
    } catch (...) {
        return -3;
    }
}

SDK_API int sdk_engine_load_seeder(void* h, const char* path) {
    if (!h || !path) return -3;
    try {
        return ((DevEngine*)h)->eng.load_seeder(path);
    } catch (...) {
        return -3;
    }
}

SDK_API int sdk_engine_launch_log(void* h, int32_t* out, int cap) {
    if (!h || (cap > 0 && !out) || cap < 0) return -3;
    const auto& log = ((DevEngine*)h)->eng.launch_log;
    const int n = (int)log.size();
    for (int i = 0; i < n && i < cap; ++i) {
        out[2 * i] = log[i].first;
        out[2 * i + 1] = log[i].second;
    }
    return n;
}

SDK_API int sdk_engine_seed_stats(void* h, int32_t* out3) {
    if (!h || !out3) return -3;
    const auto& st = ((DevEngine*)h)->eng.seed_stats;
    out3[0] = st.cnn;
    out3[1] = st.cnn_ok;
    out3[2] = st.fallback;
    return 0;
}

}  // extern "C"
