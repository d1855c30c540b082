// The drop-in C ABI: the eleven semper_* entry points of the reference
// (include/semper/semper_c.h:83-197, adapters/c/semper_c.cpp) over the MI355X
// engine. Handles, the one-call-per-handle rule, the cancel window and the
// exception barrier behave as the reference's. Image bytes are raw gray,
// raw RGBA or an encoded PNG / TIFF / BMP / PNM / JPEG file (codec.hpp);
// WebP does not decode here and returns SEMPER_ERR_INIT.
#include <atomic>
#include <cmath>
#include <cstring>
#include <exception>
#include <functional>
#include <memory>
#include <vector>

#include "codec.hpp"
#include "engine.hpp"
#include "handles.hpp"
#include "incremental.hpp"
#include "semper_c.h"

using sdic::Cancel;
using sdic::Engine;
using sdic::Params;


namespace {

// One call at a time per handle (adapters/c/semper_c.cpp:44-56).
class HandleCall {
public:
    explicit HandleCall(std::atomic<bool>& b) : busy_(b), owned_(!b.exchange(true, std::memory_order_acquire)) {}
    ~HandleCall() {
        if (owned_) busy_.store(false, std::memory_order_release);
    }
    bool owned() const { return owned_; }

private:
    std::atomic<bool>& busy_;
    bool owned_;
};

struct CancelScope {
    Cancel& c;
    explicit CancelScope(Cancel& x) : c(x) { c.begin(); }
    ~CancelScope() { c.end(); }
};

template <typename R, typename F>
R guarded(R on_error, F&& body) noexcept {
    try {
        return body();
    } catch (...) {
        return on_error;
    }
}

// Image bytes to the engine's 8-bit gray plane: raw gray / RGBA by length,
// else an encoded file (io::decode_gray, src/io/image_codec.cpp:129-161).
bool decode_image(const uint8_t* data, size_t len, int ew, int eh, std::vector<uint8_t>& out, int& w, int& h) {
    return sdic::codec::decode_gray(data, len, ew, eh, out, w, h);
}

// No mask bytes is no mask; bytes that do not decode are an error
// (adapters/c/semper_c.cpp:70-79). A mask of another size is brought to the
// image's by nearest neighbour (full_field_solver.cpp:190-194,
// incremental_session.cpp:238-242).
bool decode_mask(const uint8_t* mask, size_t len, int w, int h, std::vector<uint8_t>& out, bool& have) {
    have = mask != nullptr && len != 0;
    if (!have) return true;
    std::vector<uint8_t> m;
    int mw = 0, mh = 0;
    if (!sdic::codec::decode_mask(mask, len, w, h, m, mw, mh)) return false;
    if (mw == w && mh == h) out.swap(m);
    else sdic::codec::resize_nearest(m, mw, mh, w, h, out);
    return true;
}

Params to_params(const semper_params* p) {
    Params q;
    q.rect_x = p->rect_x;
    q.rect_y = p->rect_y;
    q.rect_w = p->rect_w;
    q.rect_h = p->rect_h;
    q.step = p->step;
    q.subset = p->subset_size;
    q.strain_window = p->strain_window;
    q.use6 = p->use_6x6_interpolator != 0;
    return q;
}

std::function<void(int)> to_progress(semper_progress_cb cb, void* user) {
    std::function<void(int)> f;
    if (cb) f = [cb, user](int pct) { cb(pct, user); };
    return f;
}

}  // namespace

extern "C" {

SEMPER_C_API semper_engine* semper_create(void) {
    return guarded<semper_engine*>(nullptr, [] { return new semper_engine(); });
}

SEMPER_C_API void semper_destroy(semper_engine* eng) {
    guarded<int>(0, [&] {
        delete eng;
        return 0;
    });
}

SEMPER_C_API int semper_set_reference(semper_engine* eng, const uint8_t* img, size_t img_len, const uint8_t* mask,
                                      size_t mask_len, int expected_w, int expected_h) {
    if (eng == nullptr) return SEMPER_ERR_INIT;
    HandleCall call(eng->busy);
    if (!call.owned()) return SEMPER_ERR_INIT;
    eng->eng.reset();
    if (img == nullptr || img_len == 0) return SEMPER_ERR_INIT;
    return guarded(SEMPER_ERR_INIT, [&] {
        std::vector<uint8_t> ref, m;
        int w = 0, h = 0;
        bool have_mask = false;
        if (!decode_image(img, img_len, expected_w, expected_h, ref, w, h)) return SEMPER_ERR_INIT;
        if (!decode_mask(mask, mask_len, w, h, m, have_mask)) return SEMPER_ERR_INIT;
        return eng->eng.set_reference(ref.data(), false, w, h, have_mask ? m.data() : nullptr);
    });
}

SEMPER_C_API int semper_run(semper_engine* eng, const uint8_t* def, size_t def_len, const uint8_t* mask,
                            size_t mask_len, const semper_params* params, float* out, int out_capacity,
                            float* metrics, int metrics_len, semper_progress_cb cb, void* user) {
    if (eng == nullptr || def == nullptr || def_len == 0 || params == nullptr || out == nullptr || out_capacity < 0)
        return SEMPER_ERR_INIT;
    HandleCall call(eng->busy);
    if (!call.owned()) return SEMPER_ERR_INIT;
    return guarded(SEMPER_ERR_INIT, [&] {
        CancelScope window(eng->cancel);
        if (!eng->eng.has_reference()) return SEMPER_ERR_INIT;
        const int w = eng->eng.width(), h = eng->eng.height();
        std::vector<uint8_t> d, m;
        int dw = 0, dh = 0;
        bool have_mask = false;
        // The frame is correlated pixel for pixel with the reference: another
        // size is an error (full_field_solver.cpp:168-174).
        if (!decode_image(def, def_len, w, h, d, dw, dh) || dw != w || dh != h) return SEMPER_ERR_INIT;
        if (!decode_mask(mask, mask_len, w, h, m, have_mask)) return SEMPER_ERR_INIT;
        eng->cancel.enter();
        return eng->eng.run(d.data(), false, have_mask ? m.data() : nullptr, to_params(params), out, out_capacity,
                            metrics, metrics_len, eng->cancel, to_progress(cb, user));
    });
}

SEMPER_C_API void semper_cancel(semper_engine* eng) {
    if (eng != nullptr) eng->cancel.request();
}

SEMPER_C_API const char* semper_version(void) { return "0.6.1"; }

SEMPER_C_API semper_incremental* semper_incremental_create(void) {
    return guarded<semper_incremental*>(nullptr, [] { return new semper_incremental(); });
}

SEMPER_C_API void semper_incremental_destroy(semper_incremental* inc) {
    guarded<int>(0, [&] {
        delete inc;
        return 0;
    });
}

SEMPER_C_API int semper_incremental_init(semper_incremental* inc, const uint8_t* ref, size_t ref_len,
                                         const uint8_t* mask, size_t mask_len, int expected_w, int expected_h,
                                         const semper_params* params, int32_t update_interval) {
    if (inc == nullptr) return SEMPER_ERR_INIT;
    HandleCall call(inc->busy);
    if (!call.owned()) return SEMPER_ERR_INIT;
    inc->session.reset();
    inc->width = inc->height = 0;
    if (ref == nullptr || ref_len == 0 || params == nullptr) return SEMPER_ERR_INIT;
    return guarded(SEMPER_ERR_INIT, [&] {
        std::vector<uint8_t> r, m;
        int w = 0, h = 0;
        bool have_mask = false;
        if (!decode_image(ref, ref_len, expected_w, expected_h, r, w, h)) return SEMPER_ERR_INIT;
        if (!decode_mask(mask, mask_len, w, h, m, have_mask)) return SEMPER_ERR_INIT;
        const int rc = inc->session.init(r.data(), w, h, have_mask ? m.data() : nullptr, to_params(params),
                                         update_interval);
        inc->width = rc == 0 ? w : 0;
        inc->height = rc == 0 ? h : 0;
        return rc;
    });
}

SEMPER_C_API int semper_incremental_process(semper_incremental* inc, const uint8_t* def, size_t def_len,
                                            float* out, int out_capacity, float* metrics, int metrics_len,
                                            semper_incremental_info* info, semper_progress_cb cb, void* user) {
    if (inc == nullptr || def == nullptr || def_len == 0 || out == nullptr || out_capacity < 0)
        return SEMPER_ERR_INIT;
    HandleCall call(inc->busy);
    if (!call.owned()) return SEMPER_ERR_INIT;
    if (!inc->session.initialized()) return SEMPER_ERR_INIT;
    return guarded(SEMPER_ERR_INIT, [&] {
        CancelScope window(inc->cancel);
        std::vector<uint8_t> d;
        int dw = 0, dh = 0;
        if (!decode_image(def, def_len, inc->width, inc->height, d, dw, dh) || dw != inc->width ||
            dh != inc->height)
            return SEMPER_ERR_INIT;
        inc->cancel.enter();
        sdic::FrameInfo fi;
        const int rc = inc->session.process(d.data(), out, out_capacity, metrics, metrics_len, inc->cancel,
                                            to_progress(cb, user), &fi);
        if (info != nullptr && rc >= 0) {
            info->reference_frame = fi.reference_frame;
            info->reference_updated = fi.reference_updated ? 1 : 0;
            info->tracked_points = fi.tracked_points;
            info->frame0_points = fi.frame0_points;
            info->taylor_fallbacks = fi.taylor_fallbacks;
            info->lost_this_frame = fi.lost_this_frame;
        }
        return rc;
    });
}

SEMPER_C_API void semper_incremental_cancel(semper_incremental* inc) {
    if (inc != nullptr) inc->cancel.request();
}

}  // extern "C"
