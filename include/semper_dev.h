/*
 * Kernel-level C interface of the MI355X DIC engine (libsemper_dev.so).
 *
 * These entry points have no counterpart in the reference's C ABI: they expose
 * the pieces behind semper_run -- the subset solver, the Hessian pool, the
 * reference planes, an engine whose frames already sit in HBM, and the
 * synthetic speckle generator -- to the test suite and the benchmark. The
 * subset solver mirrors the reference's OptimizationEngine::calculate_deformation
 * (src/math/optimization_engine.cpp:55-118) over precompute_subset
 * (src/math/subset_precomputer.cpp:31-35), node by node. libsemper_dev.so also
 * exports every semper_* function of semper_c.h.
 */
#ifndef SEMPER_DEV_H
#define SEMPER_DEV_H

#include <stddef.h>
#include <stdint.h>

#define SDK_API __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

SDK_API int sdk_device_count(void);
SDK_API void* sdk_malloc(size_t bytes);
SDK_API void sdk_free(void* p);
/* Page-locked host memory (hipHostMalloc): an output buffer of semper_run /
   sdk_engine_run in it is written by a direct copy from the device, without
   the staging that pageable memory takes. NULL on failure. */
SDK_API void* sdk_host_alloc(size_t bytes);
SDK_API void sdk_host_free(void* p);
/* kind 0: host to device, 1: device to host, 2: device to device */
SDK_API int sdk_memcpy(void* dst, const void* src, size_t bytes, int kind);

/* calculate_deformation per node on host images; see devapi.cpp. */
SDK_API int sdk_solve_nodes(const float* ref, const void* def, int def_is_f32, int w, int h, int dim, int use6, int lm,
                            float tol, int n, const int32_t* cxy, const float* guess, const int32_t* modes,
                            int32_t* iout, float* fout);

/* sdk_solve_nodes for callers that cannot allocate device memory (code under
   HIP graph capture, serving loops); it has no counterpart in the reference.
   The arguments and results are sdk_solve_nodes', but every pointer is device
   memory -- ref (w x h floats), def (w x h floats or bytes), cxy (2n int32),
   guess (6n floats), modes (n int32), iout (5n int32), fout (7n floats) --
   and the call only enqueues the work on `stream` (a hipStream_t; NULL: the
   null stream): the outputs are ready when the stream's work is. It calls no
   allocator: the solver's own buffers (reference planes, per-node pool
   entries, tasks, results, the fast path's hand-off list) live in one fixed
   device array of SDK_NOALLOC_SCRATCH_BYTES per GPU, held until the process
   ends. w x h <= SDK_NOALLOC_MAX_PIXELS and n <= SDK_NOALLOC_MAX_NODES, else
   SDK_ERR_NOALLOC_CAPACITY and nothing is enqueued. Calls that overlap on one
   GPU share that array and race: serialise them (one stream, or one call at a
   time). 0, -3 (bad arguments or a failed launch) or
   SDK_ERR_NOALLOC_CAPACITY. */
#define SDK_NOALLOC_MAX_PIXELS (1 << 20)
#define SDK_NOALLOC_MAX_NODES 16384
#define SDK_NOALLOC_SCRATCH_BYTES 17105152
#define SDK_ERR_NOALLOC_CAPACITY (-7)
SDK_API int sdk_solve_nodes_noalloc(const float* ref, const void* def, int def_is_f32, int w, int h, int dim, int use6,
                                    int lm, float tol, int n, const int32_t* cxy, const float* guess,
                                    const int32_t* modes, int32_t* iout, float* fout, void* stream);

/* compute_hessian_only per node centre (subset_precomputer.cpp:42-149). */
SDK_API int sdk_pool(const float* ref, int w, int h, int dim, int lm, int n, const int32_t* cxy, int32_t* valid,
                     float* stats2, float* hstep);

/* Image + prepare_data(false) + Ghost Wall (image_processor.cpp:9-64,
   reference_cache.cpp:18-35). */
SDK_API int sdk_ref_planes(const uint8_t* img, const uint8_t* mask, int w, int h, float* I, float* gx, float* gy);

/* An engine over frames in host or device memory (on_device != 0). */
SDK_API void* sdk_engine_create(void);
SDK_API void sdk_engine_destroy(void* h);
SDK_API int sdk_engine_set_reference(void* h, const uint8_t* img, int on_device, int w, int hh, const uint8_t* mask);
SDK_API int sdk_engine_run(void* h, const uint8_t* def, int on_device, const int32_t* params, float* out, int cap,
                           float* metrics, int metrics_len, float tol);
/* The solver's wave reductions on nw x 64 floats: every lane's float sum into
   out and integer sum of (int)in into iout (a self-check of the cross-lane
   network). */
SDK_API int sdk_wave_sum_probe(const float* in, int nw, float* out, int32_t* iout);

/* The subset solver's Keys 4x4 sampling forms (no reference counterpart;
   the interpolant is interpolate_bicubic, image_processor.cpp:188-242) at n
   points of a u8 frame (host memory): form 0 the tap weights over the FP32
   plane, 1 the row-cubic coefficient plane of flag 17, 2 the tap weights over
   the u8 plane. Every point must lie inside the solver's guard, 4 <= x < w - 5
   and 4 <= y < h - 5, and w h < 2^27; 0 or -3. */
SDK_API int sdk_interp_probe(const uint8_t* img, int w, int h, int form, int n, const float* xs, const float* ys,
                             float* out);

/* The incremental session's FP64 composition kernel (no reference
   counterpart; the sampling is sample_increment, incremental_math.cpp:61-123)
   on host arrays, for tests: an increment field (7 planes of fgw x fgh, u v ux
   uy vx vy corr, and the median test's flags, grid origin (frx, fry) and
   step) sampled at every alive node of a gw x gh grid at (rx + i step + uf,
   ry + j step + vf), the sample added to u64 / v64 in double. Per node:
   flags (bit 0 sampled, bit 1 by the first-order expansion), the new totals
   in double and rounded once, and the sample's corr (other nodes: flags 0,
   outputs unspecified). 0 or -3. */
SDK_API int sdk_compose64_probe(const float* planes, const uint8_t* med, int fgw, int fgh, int frx, int fry,
                                int fstep, const float* uf, const float* vf, const double* u64, const double* v64,
                                const uint8_t* alive, int gw, int gh, int rx, int ry, int step, uint8_t* flags,
                                double* ou64, double* ov64, float* ouf, float* ovf, float* ocorr);

/* Engine switches: flag 0 = speculative Path B first claims (default 1);
   flag 1 = `out` of sdk_engine_run is device memory (default 0);
   precision variants with no reference counterpart (default 0, a change
   rebuilds the Hessian pool on the next run): flag 2 = the Gauss-Newton step
   by FP64 Cholesky of each pool entry, flag 3 = 16-bit storage of the
   reference gradient planes (1: fp16, 2: bf16), flag 4 = J^T J of the pool on MFMA (fp16 inputs, FP32
   accumulation); flag 5 = learned anchor seeding (needs a seeder loaded by
   sdk_engine_load_seeder; -3 otherwise); flag 6 = the launch-shape tuner
   (default 1): the first runs of each (grid, subset size, interpolator, node
   count) launch Path A in each shape in turn -- 1 / 2 / 4 subsets per
   workgroup, row-major / 8 x 8-tile node order -- two runs per shape, timed
   by device events (the field does not depend on the shape), and later runs
   use the shape of least mean time; flag 7 = 1..6 forces candidate value - 1
   of the order given at sdk_engine_tuning on every Path A launch (where only
   whole grids take tiles, a listed launch takes the same workgroup width
   row-major; 0: the tuner decides); flag 8 = layered Path B with the given
   reach, 1 .. 32 (default 0: the reference's flood of 32 pops per round;
   learned seeding uses reach 16 unless set): each layer claims every open node
   within `reach` grid steps of an accepted one, seeded from the accepted node
   a breadth-first search in (corr, y, x) order reaches it from first, and
   solves them in one launch; a failure beyond one step leaves the node open
   for a closer claim; flag 9 = graph capture of the fixed per-frame kernel
   chains (default 1): the end of a run -- valid flags, the five median
   passes, VSG, the row-major pack, the count download -- is captured the
   second time it is met with the same grid, parameters and buffers and
   replayed from then on; flag 11 = the learned seeder's inference on MFMA
   (default 1: fp16 activations and weights, FP32 accumulation, one
   workgroup per patch pair; 0: the FP32 forward pass it is trained with);
   flag 12 = keep a per-node report of the next runs (default 0; see
   sdk_engine_node_report); flag 13 = the solver samples an FP32 copy of the
   deformed frame made once per run (default 1; the same field bit for bit)
   rather than unpacking the u8 frame in every gather; flag 14 = a CNN-seeded
   anchor whose IC-GN fails takes the Simplex rescue before the plain anchor
   solve (default 1; 0: the plain solve at once); flag 16 = the global-motion
   search that seeds the anchor lattice (default 1; 0: no search, the lattice
   and Path C start from a zero shift, as the reference's do when its feature
   match finds no model); flag 17 = Keys 4x4 from a row-cubic coefficient
   plane (default 0): each run makes, from the u8 frame, 16 bytes per pixel
   (the Keys cubics of the cell to the pixel's right on its row and the next,
   four fp16 each, exact for u8 taps), and the fast path's interior samples
   read two such entries per sample and evaluate each row by Horner, instead
   of one load and four weighted taps per stencil row (the same interpolant,
   other roundings; Keys 6x6, flag 13 off and frames of 2^27 pixels or more
   keep the direct form); flag 18 = the layered Path B's frontier (sources,
   breadth-first search, claims, updates) on the device (default 1; for reach
   <= 16 on one subset size; 0: on the host; the field is the same either
   way); flag 19 = under learned seeding (flag 5), the CNN lattice on every
   frame (default 0: only where the global model finds more than a
   near-translation or the plain lattice's anchors do not converge). */
SDK_API int sdk_engine_set_flag(void* h, int flag, int value);

/* Phase timings of the last run, 14 doubles: total, prep, seeding, prepass,
   mesh, Path A, Path B, strain (host ms); Path A solver launch (device ms),
   Path A items, Path A IC-GN iterations, Path B speculative solves, further
   Path B launches, subsets those launches solved. */
SDK_API int sdk_engine_timings(void* h, double* out14);

/* The subset-solver launches of the last run, in order: for each, the work
   items and how many of them the fast path handed to the full solver
   (masked iterations, coarse searches, Simplex rescues). Writes up to cap
   (items, handed) pairs to out; returns the launch count, or -3. */
SDK_API int sdk_engine_launch_log(void* h, int32_t* out, int cap);

/* Mixed subset sizes (no reference counterpart): the subset size of each node
   of a gw x gh grid (row-major, odd, 3 .. 63) for the engine's runs on a grid
   of that shape; dims NULL clears the map. Each node is solved with its own
   size (scan border, Hessian pool, IC-GN); one solver launch per size.
   0 or -3. */
SDK_API int sdk_engine_set_subset_map(void* h, int gw, int gh, const int32_t* dims);

/* Mixed subset sizes by image region (no reference counterpart): an rw x rh
   map (row-major, odd sizes 3 .. 63) over the frame; a node at pixel (x, y)
   takes the size of region (x rw / w, y rh / h), on any grid. A per-node map
   of the run's grid shape (sdk_engine_set_subset_map) takes precedence; rw =
   rh = 0 clears. The session form applies to every solve of an incremental
   session created by semper_incremental_create, on each adopted reference
   too, and is kept across semper_incremental_init. 0 or -3. */
SDK_API int sdk_engine_set_region_map(void* h, int rw, int rh, const int32_t* dims);
struct semper_incremental;
SDK_API int sdk_incremental_set_region_map(struct semper_incremental* inc, int rw, int rh, const int32_t* dims);

/* Per-node provenance of the last run with flag 12 on (the reference's
   debug export, full_field_debug_export.cpp, in a form of its own), row-major
   over its grid: tag = the mesh tag of the node's guess (0 none, 1 inside a
   lattice triangle, 2 extrapolated, 3 anchor, 4 global model), path = 0
   skipped by the scan, 1 unsolved, 2 anchor, 3 Path A, 4 a later path. The
   node count, or -3 (no report, or n too small). */
SDK_API int sdk_engine_node_report(void* h, uint8_t* tag, uint8_t* path, int n);
/* The node planes of the same report: u, v, ux, uy, vx, vy, corr, plane after
   plane (7 x node count floats), at the end of the run (which 0) or after
   Path A (which 1); which 2: the six guess planes (u, v, ux, uy, vx, vy) of
   the mesh guess field Path A started from (6 x node count floats). The float
   count, or -3. */
SDK_API int sdk_engine_node_planes(void* h, int which, float* out, int n);

/* The global-motion model the last run seeded its anchor lattice with (the
   replacement of the reference's AKAZE homography, full_field_anchors.cpp:
   204-279): out12 = {locked, full, a0, a1, a2, a3, a4, a5, cx, cy, patches,
   pass}, with u = a0 + a1 (x - cx) + a2 (y - cy) and v = a3 + a4 (x - cx) +
   a5 (y - cy) at reference pixel (x, y); full = an ROI corner moves more than
   4 px beyond the centre's translation; pass = the search that locked it (1
   translation, 2 the wider translation search, 3 the rotation-invariant
   search; 0 none). 0 or -3. */
SDK_API int sdk_engine_global_model(void* h, double* out12);

/* Graph captures and replays of an engine so far: stats2 = {captures,
   replays}. 0 or -3. */
SDK_API int sdk_engine_graph_stats(void* h, int32_t* stats2);
/* The session forms: flag 9 (graph capture, default 1; applies from the next
   semper_incremental_init, which also runs the session's strain of the total
   field on the frame-0 grid through a captured chain), flag 10 (FP64
   composition of the total displacement, default 0; from the next init: the
   totals carried across reference updates are kept in double, each increment
   added in double and the reported total rounded once, where the reference
   adds increments to a float total), and the captures and replays of the
   session's engine. 0 or -3. */
SDK_API int sdk_incremental_set_flag(struct semper_incremental* inc, int flag, int value);
SDK_API int sdk_incremental_graph_stats(struct semper_incremental* inc, int32_t* stats2);
/* Host milliseconds of the session's last frame, 8 doubles: total, the
   engine's run, the composition of the totals (sampling, re-anchoring,
   strain of the total field), the reference update, and within the update
   the field fill, the mask warp, the warped mask's bounds and the adoption
   of the frame as the new reference. 0 or -3. */
SDK_API int sdk_incremental_timings(struct semper_incremental* inc, double* t8);
/* The session engine's sdk_engine_timings for its last run (the frame's
   solve against the current reference). 0, or -3 before init. */
SDK_API int sdk_incremental_engine_timings(struct semper_incremental* inc, double* out14);
/* The last frame's re-anchoring: out2 = {stored subsets matched on the frame
   (nodes that coasted past a quarter subset and came back), of those
   accepted}. 0 or -3. */
SDK_API int sdk_incremental_anchor_stats(struct semper_incremental* inc, int32_t* out2);

/* The tuner's latest timing: chosen2 = {subsets per workgroup, tile side (0:
   row-major)} of least mean time so far, ms6 = the mean Path A launch time
   per candidate in the order (1, 0) (2, 0) (4, 0) (1, 8) (2, 8) (4, 8), -1
   where not timed. 0 once the choice is settled, 1 while still timing, -3
   when nothing has been timed. */
SDK_API int sdk_engine_tuning(void* h, int32_t* chosen2, double* ms6);

/* Synthetic speckle pair (synth.hpp FieldSpec kinds). */
SDK_API int sdk_synth_pair(uint64_t seed, int w, int h, int kind, const double* p, int np, double cx, double cy,
                           float noise_sigma, uint8_t* ref, uint8_t* def, int on_device);
/* One deformed frame of a sequence: the speckle of `seed` under the field, with
   its own noise drawn from noise_seed; w x h bytes to def (host or device). */
SDK_API int sdk_synth_frame(uint64_t seed, uint64_t noise_seed, int w, int h, int kind, const double* p, int np,
                            double cx, double cy, float noise_sigma, uint8_t* def, int on_device);

/* The boundary's image decode (host only, no GPU).
   sdk_decode: io::decode_gray (mask = 0) or io::decode_mask (mask = 1)
   (src/io/image_codec.cpp:129-161,254-262): writes the decoded size to *w, *h
   and, when out is not NULL and cap >= w * h, the plane. Returns 0, or -3 when
   nothing decodes (or cap is short).
   sdk_image_dimensions: io::image_dimensions (image_codec.cpp:265-287).
   sdk_to_gray8 / sdk_to_mask8: io::to_gray8 / to_mask8 (image_codec.cpp:221-252)
   of a w x h image of cn = 1 or 3 interleaved channels (R, G, B) of
   depth 0 = u8, 1 = u16, 2 = s16, 3 = s32, 4 = f32, 5 = f64; w * h bytes to
   out. Returns 0 or -3. */
SDK_API int sdk_decode(const uint8_t* data, size_t len, int expected_w, int expected_h, int mask, uint8_t* out,
                       size_t cap, int32_t* w, int32_t* h);
SDK_API void sdk_image_dimensions(const uint8_t* data, size_t len, int32_t* w, int32_t* h);
SDK_API int sdk_to_gray8(const void* px, int w, int h, int cn, int depth, uint8_t* out);
SDK_API int sdk_to_mask8(const void* px, int w, int h, int cn, int depth, uint8_t* out);

/* The learned seeder (no reference counterpart): a CNN from a level-2 patch
   pair (2 x 64 x 64, each plane normalised) to the scaled warp (u, v) / 16,
   (ux, uy, vx, vy) / 0.1 of the patch centre; see csrc/seeder.hpp.
   sdk_seeder_param_count: the flat parameter count.
   sdk_seeder_grad / sdk_seeder_forward: one forward (and backward) pass on the
   GPU with the given parameters on a host batch; the loss is the mean squared
   error over B x 6, grads its gradient.
   sdk_seeder_train: renders a 2048 x 2048 speckle reference from `seed`,
   trains from a He initialisation (steps x batch pairs, Adam, cosine decay
   from lr), writes the weights to out_path; returns the count of mean losses
   (one per 100 steps) written to losses (at most nloss), or -3.
   sdk_engine_load_seeder: weights for an engine (0 or -3).
   sdk_engine_seed_stats: anchors given CNN guesses, those whose refined
   solve held, those solved again the plain way, of the last run. */
SDK_API int sdk_seeder_param_count(void);
SDK_API int sdk_seeder_grad(const float* params, const float* x, const float* t, int batch, float* grads,
                            float* loss);
SDK_API int sdk_seeder_forward(const float* params, const float* x, int batch, float* y);
/* The same batch through the inference kernel the engine uses: fp16
   activations and weights on MFMA (v_mfma_f32_16x16x16f16) with FP32
   accumulation. */
SDK_API int sdk_seeder_forward_mfma(const float* params, const float* x, int batch, float* y);
SDK_API int sdk_seeder_train(uint64_t seed, int steps, int batch, float lr, const char* out_path, float* losses,
                             int nloss);
SDK_API int sdk_engine_load_seeder(void* h, const char* path);
SDK_API int sdk_engine_seed_stats(void* h, int32_t* out3);

#ifdef __cplusplus
}
#endif

#endif /* SEMPER_DEV_H */
