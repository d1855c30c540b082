// CPU ORACLE — test infrastructure only (see oracle.hpp).
// C entry points the tests and the benchmark's CPU-baseline leg load with ctypes.
#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "oracle.hpp"
#include "seeding.hpp"
#include "synth.hpp"

using namespace orc;

extern "C" {

// ---- synthetic images ------------------------------------------------------

// tests/framework/synthetic.h: reference (aff == null) or deformed image.
void orc_ref_speckle(unsigned seed, int w, int h, int blobs, const float* aff, float* out) {
    RefSpeckle f(seed, w, h, blobs);
    f.render(w, h, aff, out);
}

void orc_gray8(const float* in, int w, int h, float noise_sigma, unsigned seed, uint8_t* out) {
    gray8(in, w, h, noise_sigma, seed, out);
}

// Benchmark speckle S at arbitrary points (xs, ys), unnormalised.
void orc_blob_eval(uint64_t seed, int w, int h, float density, float pad, int n, const float* xs,
                   const float* ys, float* out) {
    BlobField f(seed, w, h, density, pad);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) out[i] = f.eval(xs[i], ys[i]);
}

float orc_hash_u01(uint64_t seed, uint64_t i) { return hash_u01(seed, i); }

// ---- seeding front end -------------------------------------------------------

// The global-motion model run_full_field seeds the anchor lattice with, on a
// u8 pair and rect (x, y, w, h): out12 = {locked, full, a0..a5, cx, cy,
// patches, 0}, u = a0 + a1 (x - cx) + a2 (y - cy), v = a3 + a4 (.) + a5 (.).
void orc_global_motion(const uint8_t* ref, const uint8_t* def, int w, int h, const int32_t* rect, double* out12) {
    const int L = pyramid_levels(rect[2], rect[3]);
    std::vector<float> rf(ref, ref + (size_t)w * h), df(def, def + (size_t)w * h);
    Pyramid pr, pd;
    build_pyramid(pr, rf.data(), w, h, L);
    build_pyramid(pd, df.data(), w, h, L);
    const GlobalModel gm = global_motion(pr, pd, rect[0], rect[1], rect[2], rect[3]);
    const double v[12] = {gm.locked ? 1.0 : 0.0, gm.full ? 1.0 : 0.0, gm.a[0], gm.a[1], gm.a[2], gm.a[3], gm.a[4],
                          gm.a[5], gm.cx, gm.cy, (double)gm.patches, 0.0};
    std::copy(v, v + 12, out12);
}

// ---- subset level ------------------------------------------------------------

// Golden corpus (tests/integration/test_golden_corpus.cpp:54-127): 512x512,
// seed 42 / 400 blobs, d = 27, grid 60..452 step 22, translate and affine
// scenarios, zero guess, INIT_NO_SIMPLEX, bare engine (no LM). Writes up to cap
// points as {x, y, status} ints and {u, v, ux, uy, vx, vy, corr} floats.
int orc_golden_corpus(int use6, int cap, int32_t* ixy_status, float* vals) {
    const int W = 512, H = 512, D = 27;
    const float sc[2][6] = {{3.2f, -1.7f, 0.0f, 0.0f, 0.0f, 0.0f},
                            {2.1f, 1.4f, 0.015f, -0.008f, 0.006f, 0.012f}};
    int n = 0;
    std::vector<float> rbuf((size_t)W * H), dbuf((size_t)W * H);
    for (int s = 0; s < 2; ++s) {
        RefSpeckle f(42, W, H, 400);
        const float aff[8] = {sc[s][0], sc[s][1], sc[s][2], sc[s][3],
                              sc[s][4], sc[s][5], W / 2.0f, H / 2.0f};
        f.render(W, H, nullptr, rbuf.data());
        f.render(W, H, aff, dbuf.data());
        Image ref, def;
        image_from_f32(ref, W, H, rbuf.data());
        image_from_f32(def, W, H, dbuf.data());
        for (int y = 60; y <= 452; y += 22)
            for (int x = 60; x <= 452; x += 22) {
                Subset sub;
                precompute_subset(sub, ref, x, y, D);
                if (!sub.init) continue;
                Engine e;
                e.use6 = use6 != 0;
                const float g[6] = {0, 0, 0, 0, 0, 0};
                const Result r = e.calculate_deformation(sub, def, g, NO_SIMPLEX);
                if (n < cap) {
                    ixy_status[3 * n] = x;
                    ixy_status[3 * n + 1] = y;
                    ixy_status[3 * n + 2] = r.status;
                    const float v[7] = {r.u, r.v, r.ux, r.uy, r.vx, r.vy, r.corr};
                    for (int k = 0; k < 7; ++k) vals[7 * n + k] = v[k];
                }
                ++n;
            }
    }
    return n;
}

// Solve a list of nodes on float images (the same entry the GPU node solver
// offers): per node centre (cx, cy), a 6-DOF guess and a mode. lm selects the
// full-field engine configuration (LM on the translation DOFs). Out per node:
// ints {init, status, iters, trigger_iters, simplex} and floats
// {u, v, ux, uy, vx, vy, corr}.
void orc_solve_nodes(const float* ref_img, const float* def_img, int w, int h, int dim, int use6,
                     int lm, float conv_tol, int n, const int32_t* cxy, const float* guess,
                     const int32_t* modes, int32_t* iout, float* fout) {
    Image ref, def;
    image_from_f32(ref, w, h, ref_img);
    image_from_f32(def, w, h, def_img);
#pragma omp parallel
    {
        Engine e;
        e.use6 = use6 != 0;
        e.lm = lm != 0;
        e.lm_alpha = lm ? kLmAlpha : 0.0f;
        e.conv_tol = conv_tol;
        Subset sub;
#pragma omp for schedule(dynamic, 8)
        for (int i = 0; i < n; ++i) {
            precompute_subset(sub, ref, cxy[2 * i], cxy[2 * i + 1], dim);
            int32_t* io = iout + 5 * i;
            float* fo = fout + 7 * i;
            io[0] = sub.init ? 1 : 0;
            if (!sub.init) {
                io[1] = 1;
                io[2] = io[3] = io[4] = 0;
                for (int k = 0; k < 7; ++k) fo[k] = 0.0f;
                continue;
            }
            const int before = e.simplex_calls;
            Result r = e.calculate_deformation(sub, def, guess + 6 * i, (Mode)modes[i]);
            // full-field glue: reject_if_ghosted (full_field_internal.hpp:347-353)
            if (lm && (float)r.invalid_ref / (float)(dim * dim) > kGhostRejectFraction) {
                r.corr = 2.0f;
                r.status = 1;
            }
            io[1] = r.status;
            io[2] = r.iters;
            io[3] = r.trigger_iters;
            io[4] = e.simplex_calls - before;
            const float v[7] = {r.u, r.v, r.ux, r.uy, r.vx, r.vy, r.corr};
            for (int k = 0; k < 7; ++k) fo[k] = v[k];
        }
    }
}

// Hessian pool entries (compute_hessian_only) for a node list: valid, mean,
// std, H[36], Hinv[36] per node.
void orc_pool(const float* ref_img, int w, int h, int dim, int n, const int32_t* cxy,
              int32_t* valid, float* stats2, float* H, float* Hinv) {
    Image ref;
    image_from_f32(ref, w, h, ref_img);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const PoolEntry e = compute_hessian_only(ref, cxy[2 * i], cxy[2 * i + 1], dim);
        valid[i] = e.valid;
        stats2[2 * i] = e.mean;
        stats2[2 * i + 1] = e.stdv;
        std::copy(e.H, e.H + 36, H + 36 * i);
        std::copy(e.Hinv, e.Hinv + 36, Hinv + 36 * i);
    }
}

// Gradients of a float image (prepare_data(false)).
void orc_gradients(const float* img, int w, int h, float* gx, float* gy) {
    Image m;
    image_from_f32(m, w, h, img);
    std::copy(m.gx.begin(), m.gx.end(), gx);
    std::copy(m.gy.begin(), m.gy.end(), gy);
}

// Interpolate a float image at points; kind 0 bicubic, 1 keys 6x6, 2 bilinear.
void orc_interp(const float* img, int w, int h, int kind, int n, const float* xs, const float* ys,
                float* out) {
    Image m;
    m.w = w;
    m.h = h;
    m.I.assign(img, img + (size_t)w * h);
    for (int i = 0; i < n; ++i)
        out[i] = kind == 0 ? interp_bicubic(m, xs[i], ys[i])
               : kind == 1 ? interp_keys6(m, xs[i], ys[i])
                           : interp_bilinear(m, xs[i], ys[i]);
}

// FP64 IC-GN of the same model from per-node starts p0 (6 doubles each).
// Out: status, iters per node; p (6 doubles) and corr.
void orc_icgn_fp64(const float* ref_img, const float* def_img, int w, int h, int dim, int use6,
                   int n, const int32_t* cxy, const double* p0, double tol, int max_iter,
                   int32_t* status_iters, double* pout, double* corr) {
    Image ref, def;
    ref.w = def.w = w;
    ref.h = def.h = h;
    ref.I.assign(ref_img, ref_img + (size_t)w * h);
    def.I.assign(def_img, def_img + (size_t)w * h);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < n; ++i) {
        const Result64 r =
            icgn_fp64(ref, def, cxy[2 * i], cxy[2 * i + 1], dim, use6 != 0, p0 + 6 * i, tol, max_iter);
        status_iters[2 * i] = r.status;
        status_iters[2 * i + 1] = r.iters;
        std::copy(r.p, r.p + 6, pout + 6 * i);
        corr[i] = r.corr;
    }
}

// ---- strain ----------------------------------------------------------------

int orc_vsg_window_node_count(int step, int window) { return vsg_window_node_count(step, window); }

// Median test then VSG on a node grid, as pack_full_field_output does before
// packing. valid is updated in place; returns the number rejected.
int orc_outliers_and_vsg(int gw, int gh, int step, int window, float* u, float* v,
                         uint8_t* valid, float* exx, float* eyy, float* exy) {
    DispField d;
    d.w = gw;
    d.h = gh;
    d.step = step;
    d.u.assign(u, u + (size_t)gw * gh);
    d.v.assign(v, v + (size_t)gw * gh);
    d.valid.assign(valid, valid + (size_t)gw * gh);
    const int rej = reject_displacement_outliers(d);
    const StrainField s = compute_vsg_strain(d, window);
    std::copy(d.valid.begin(), d.valid.end(), valid);
    std::copy(s.exx.begin(), s.exx.end(), exx);
    std::copy(s.eyy.begin(), s.eyy.end(), eyy);
    std::copy(s.exy.begin(), s.exy.end(), exy);
    return rej;
}

}  // extern "C"

// ---- full field (the CPU restatement of semper_run) ---------------------------

extern "C" {

struct orc_engine {
    RefCache cache;
};

orc_engine* orc_engine_create(void) { return new orc_engine(); }
void orc_engine_destroy(orc_engine* e) { delete e; }

int orc_engine_set_reference(orc_engine* e, const uint8_t* img, int w, int h, const uint8_t* mask) {
    if (!e || !img || w <= 0 || h <= 0) return kErrInit;
    set_reference(e->cache, w, h, img, mask);
    return 0;
}

// params: rect_x, rect_y, rect_w, rect_h, step, subset, strain_window, use6.
// threads 0 = all cores; conv_tol the IC-GN stop rule (1e-3 in the reference).
int orc_engine_run(orc_engine* e, const uint8_t* def, const uint8_t* mask, const int32_t* params,
                   float* out, int out_capacity, float* metrics, int metrics_len, int threads,
                   float conv_tol) {
    if (!e || !params) return kErrInit;
    Params p;
    p.rect_x = params[0];
    p.rect_y = params[1];
    p.rect_w = params[2];
    p.rect_h = params[3];
    p.step = params[4];
    p.subset = params[5];
    p.strain_window = params[6];
    p.use6 = params[7] != 0;
    RunOptions o;
    o.threads = threads;
    o.conv_tol = conv_tol;
    try {
        return run_full_field(e->cache, def, mask, p, out, out_capacity, metrics, metrics_len, o);
    } catch (...) {
        return kErrInit;
    }
}

// orc_engine_run that also writes the run's node planes: 7 planes of gw x gh
// floats (u, v, ux, uy, vx, vy, corr), corr -1 where a node holds no accepted
// result; gw = rect_w / step, gh = rect_h / step of the validated params
// (written to gwh[0..1]). planes must hold 7 x gw x gh floats (plane_cap).
int orc_engine_run_planes(orc_engine* e, const uint8_t* def, const uint8_t* mask, const int32_t* params,
                          float* out, int out_capacity, float* metrics, int metrics_len, int threads,
                          float conv_tol, float* planes, int plane_cap, int32_t* gwh) {
    if (!e || !params || !planes || !gwh) return kErrInit;
    Params p;
    p.rect_x = params[0];
    p.rect_y = params[1];
    p.rect_w = params[2];
    p.rect_h = params[3];
    p.step = params[4];
    p.subset = params[5];
    p.strain_window = params[6];
    p.use6 = params[7] != 0;
    RunOptions o;
    o.threads = threads;
    o.conv_tol = conv_tol;
    Capture cap;
    o.capture = &cap;
    int rc;
    try {
        rc = run_full_field(e->cache, def, mask, p, out, out_capacity, metrics, metrics_len, o);
    } catch (...) {
        return kErrInit;
    }
    if (rc < 0) return rc;
    const size_t n = (size_t)cap.gw * cap.gh;
    gwh[0] = cap.gw;
    gwh[1] = cap.gh;
    if ((size_t)plane_cap < 7 * n) return kErrInit;
    const std::vector<float>* pl[7] = {&cap.u, &cap.v, &cap.ux, &cap.uy, &cap.vx, &cap.vy, &cap.corr};
    for (int k = 0; k < 7; ++k) std::copy(pl[k]->begin(), pl[k]->end(), planes + k * n);
    return rc;
}

// IncrementalSession: params as orc_engine_run; info receives 6 ints
// (reference_frame, reference_updated, tracked_points, frame0_points,
// taylor_fallbacks, lost_this_frame).
void* orc_inc_create(void) { return inc_create(); }
void orc_inc_destroy(void* s) { inc_destroy((IncSession*)s); }

int orc_inc_init(void* s, const uint8_t* ref, int w, int h, const uint8_t* mask, const int32_t* params, int interval,
                 int threads, float conv_tol) {
    if (!s || !params) return kErrInit;
    Params p;
    p.rect_x = params[0];
    p.rect_y = params[1];
    p.rect_w = params[2];
    p.rect_h = params[3];
    p.step = params[4];
    p.subset = params[5];
    p.strain_window = params[6];
    p.use6 = params[7] != 0;
    RunOptions o;
    o.threads = threads;
    o.conv_tol = conv_tol;
    try {
        return inc_init((IncSession*)s, ref, w, h, mask, p, interval, o);
    } catch (...) {
        return kErrInit;
    }
}

int orc_inc_process(void* s, const uint8_t* def, float* out, int out_capacity, float* metrics, int metrics_len,
                    int32_t* info) {
    if (!s) return kErrInit;
    try {
        IncInfo fi;
        const int rc = inc_process((IncSession*)s, def, out, out_capacity, metrics, metrics_len, &fi);
        if (info && rc >= 0) {
            info[0] = fi.reference_frame;
            info[1] = fi.reference_updated;
            info[2] = fi.tracked_points;
            info[3] = fi.frame0_points;
            info[4] = fi.taylor_fallbacks;
            info[5] = fi.lost_this_frame;
        }
        return rc;
    } catch (...) {
        return kErrInit;
    }
}

}  // extern "C"
