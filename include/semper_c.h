/*
 * Semper DIC engine, MI355X build -- the C ABI.
 *
 * Drop-in for the reference engine's frozen C surface
 * (sempermechanics/semper-dic-engine include/semper/semper_c.h:83-197): the same
 * eleven functions, the same parameter and info structs, the packed output of
 * 8 floats per point [x, y, u, v, exx, eyy, exy, corr], the 24 metrics slots and
 * the return codes. libsemper_c.so exports these symbols and nothing else.
 *
 * Threading: one call at a time per handle; a concurrent or re-entrant call
 * (from another thread or from inside the progress callback) returns
 * SEMPER_ERR_INIT and touches nothing. semper_cancel / semper_incremental_cancel
 * may be called from any thread at any time.
 */
#ifndef SEMPER_C_H
#define SEMPER_C_H

#include <stddef.h>
#include <stdint.h>

#define SEMPER_C_API __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

/* Opaque engine: a cached reference (in HBM) plus a cancel token.
   Reference: adapters/c/semper_c.cpp:25-29. */
typedef struct semper_engine semper_engine;

/* Solve inputs, 8 x int32 (reference semper_c.h:46-52; offsets frozen). */
typedef struct semper_params {
    int32_t rect_x, rect_y, rect_w, rect_h; /* ROI in pixels, clipped to the image on the lattice */
    int32_t step;                           /* grid spacing */
    int32_t subset_size;                    /* odd, >= 3 (this build: <= 63) */
    int32_t strain_window;                  /* VSG window, pixels */
    int32_t use_6x6_interpolator;           /* 0: Keys 4x4 bicubic, 1: Keys 6x6 */
} semper_params;

/* Return codes (reference semper_c.h:55-64). >= 0 is a point count. */
#define SEMPER_ERR_ROI           (-2)
#define SEMPER_ERR_INIT          (-3)
#define SEMPER_ERR_STRAIN_WINDOW (-4)
#define SEMPER_ERR_SEED          (-5)
#define SEMPER_ERR_CANCELLED     (-99)

/* Progress in percent, 10..90 from a solver thread, 100 on the calling thread
   on success (reference semper_c.h:67). */
typedef void (*semper_progress_cb)(int percentage, void* user);

#define SEMPER_FLOATS_PER_POINT 8
#define SEMPER_METRICS_LEN 24

/* Reference semper_c.cpp:97-103. NULL on allocation failure. */
SEMPER_C_API semper_engine* semper_create(void);
SEMPER_C_API void semper_destroy(semper_engine* eng);

/* Caches the reference (reference semper_c.cpp:105-124): raw 8-bit gray
   (img_len == w*h) or RGBA (4*w*h) of expected_w x expected_h, or an encoded
   PNG, TIFF, BMP, PNM or JPEG file (decode_any, image_codec.cpp:129-161); mask
   of the same forms or NULL/0 (a pixel below 128 is excluded). The old
   reference is released first. 0, or SEMPER_ERR_INIT (including bytes that
   do not decode, such as WebP). */
SEMPER_C_API int semper_set_reference(semper_engine* eng, const uint8_t* img, size_t img_len,
                                      const uint8_t* mask, size_t mask_len, int expected_w, int expected_h);

/* Full-field solve of one deformed frame against the cached reference
   (reference semper_c.cpp:126-153). Writes up to out_capacity floats,
   8 per point, row-major over the grid; metrics (may be NULL) receive the
   largest allowed prefix of the 24 slots when metrics_len >= 16. */
SEMPER_C_API int semper_run(semper_engine* eng, const uint8_t* def, size_t def_len, const uint8_t* mask,
                            size_t mask_len, const semper_params* params, float* out, int out_capacity,
                            float* metrics, int metrics_len, semper_progress_cb cb, void* user);

/* Reference semper_c.cpp:155-157. Thread-safe, lock-free. */
SEMPER_C_API void semper_cancel(semper_engine* eng);

/* "MAJOR.MINOR.PATCH" of the engine surface this build implements. */
SEMPER_C_API const char* semper_version(void);

/* Incremental (updated-reference) sessions (reference semper_c.h:120-197). */
typedef struct semper_incremental semper_incremental;

/* Per-frame bookkeeping, 6 x int32 (reference semper_c.h:155-162). */
typedef struct semper_incremental_info {
    int32_t reference_frame;
    int32_t reference_updated;
    int32_t tracked_points;
    int32_t frame0_points;
    int32_t taylor_fallbacks;
    int32_t lost_this_frame;
} semper_incremental_info;

SEMPER_C_API semper_incremental* semper_incremental_create(void);
SEMPER_C_API void semper_incremental_destroy(semper_incremental* inc);

/* Reference semper_c.cpp:170-197: 0, SEMPER_ERR_ROI, SEMPER_ERR_INIT
   (decode failure, update_interval < 0) or SEMPER_ERR_STRAIN_WINDOW. */
SEMPER_C_API int semper_incremental_init(semper_incremental* inc, const uint8_t* ref, size_t ref_len,
                                         const uint8_t* mask, size_t mask_len, int expected_w, int expected_h,
                                         const semper_params* params, int32_t update_interval);

/* Reference semper_c.cpp:199-230: total displacement from the first image at
   its grid points, packed as semper_run. */
SEMPER_C_API int semper_incremental_process(semper_incremental* inc, const uint8_t* def, size_t def_len,
                                            float* out, int out_capacity, float* metrics, int metrics_len,
                                            semper_incremental_info* info, semper_progress_cb cb, void* user);

SEMPER_C_API void semper_incremental_cancel(semper_incremental* inc);

#ifdef __cplusplus
}
#endif

#endif /* SEMPER_C_H */
