/*
 * gsplat_ext.h -- additive entry points of the MI355X-native renderer.
 *
 * The reference v0.1 C ABI (gsplat.h) has no readback, no in-memory scene
 * load, no per-stage timing and no batching; the reference lists in-memory
 * scene loading as outside the v0.1 contract (README.md:167-170). Everything
 * here is additive: no struct of gsplat.h changes layout and no existing
 * symbol changes behaviour. Each declaration names the reference function
 * whose behaviour it exposes.
 *
 * Conventions are those of gsplat.h: int32_t returns are GsplatErrorCode,
 * gsplat_last_error_message() carries the detail, NULL handles or outputs
 * return GSPLAT_ERROR_INVALID_ARGUMENT.
 */
#ifndef GSPLAT_FFI_C_GSPLAT_EXT_H
#define GSPLAT_FFI_C_GSPLAT_EXT_H

#include <stddef.h>
#include <stdint.h>

#include "gsplat.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- scenes */

/*
 * Borrowed view of scene buffers in host memory, field for field the
 * reference's SceneBuffers (crates/gsplat-core/src/lib.rs:201-213):
 *   positions      count*3 floats, xyz
 *   opacity        count logits (alpha = sigmoid(logit))
 *   scale_xyz      count*3 log-scales
 *   rotation_xyzw  count*4 quaternions (need not be unit)
 *   color_dc       count*3 DC SH coefficients
 *   sh_rest        count*3*((sh_degree+1)^2-1) floats, per splat channel-major
 *                  [R c1..cK, G c1..cK, B c1..cK]; NULL iff sh_degree == 0
 * Validation follows SceneBuffers::validate (core/src/lib.rs:224-276):
 * non-finite values -> PARSE_FAILED, sh_degree > 4 -> UNSUPPORTED.
 */
typedef struct GsplatSceneView {
  uint32_t count;
  uint32_t sh_degree;
  const float *positions;
  const float *opacity;
  const float *scale_xyz;
  const float *rotation_xyzw;
  const float *color_dc;
  const float *sh_rest;
} GsplatSceneView;

/* Replaces Renderer::load_scene (render-wgpu/src/lib.rs:488-497) for a scene
 * already in memory; the PLY path is gsplat_context_load_scene_path. */
int32_t gsplat_context_load_scene_memory(GsplatContext *ctx, const GsplatSceneView *scene);

/* Renderer::set_size (render-wgpu/src/lib.rs:339-357): 0 -> INVALID_ARGUMENT,
 * above the 2D target limit (16384) -> UNSUPPORTED. */
int32_t gsplat_context_set_size(GsplatContext *ctx, uint32_t width, uint32_t height);

/* Current camera (after gsplat_context_set_auto_camera, for instance). */
int32_t gsplat_context_get_camera(const GsplatContext *ctx, GsplatCamera *out_camera);

/* Renderer::wait_for_gpu (render-wgpu/src/lib.rs:468-486). */
int32_t gsplat_context_wait(GsplatContext *ctx);

/* Renderer::readback_rgba8 (render-wgpu/src/lib.rs:779-795, 2899-2977):
 * width*height*4 bytes, rows top first, premultiplied RGBA. `len` must be at
 * least that size. Waits for the frame. */
int32_t gsplat_context_readback_rgba8(GsplatContext *ctx, uint8_t *out, uint64_t len);

/* Same image before 8-bit quantisation: width*height*4 floats. */
int32_t gsplat_context_readback_rgba32f(GsplatContext *ctx, float *out, uint64_t len);

/* Device time per stage of the last frame (hipEvent pairs on the context's
 * stream) and the sizes that drive each stage. The four stage times are
 * filled for frames rendered with stage timing on
 * (gsplat_context_set_stage_timing / gsplat_batch_set_stage_timing, or
 * GSPLAT_STAGE_TIMES=1) and are 0 otherwise: an event at each stage boundary
 * costs the device a drain (about 6 us each). total_ms is always filled. */
typedef struct GsplatStageTimes {
  float project_ms;  /* depth key + cull + projection + SH colour      */
  float sort_ms;     /* stable descending depth sort                   */
  float bin_ms;      /* tile counts, scan, pair emission, tile sort    */
  float raster_ms;   /* LDS-staged front-to-back tile blend            */
  float total_ms;    /* first event to last event                      */
  uint32_t splat_count;    /* N                                        */
  uint32_t visible_count;  /* V: near <= depth <= far                  */
  uint32_t tiles_x;
  uint32_t tiles_y;
  uint64_t tile_pairs;     /* M: (tile, splat) pairs                   */
  uint64_t blend_evals;    /* E: (pixel, splat) evaluations passing the
                            * coverage test before the pixel's early stop;
                            * counted only when GSPLAT_COUNT_EVALS=1 (a
                            * counting raster variant), else 0           */
  uint64_t blend_tests;    /* coverage tests the raster made (counting
                            * variant only, else 0)                      */
  uint64_t bin_pairs;      /* (binning tile, splat) pairs emitted and
                            * sorted; binning tiles are 16 << bin_shift
                            * pixels square                             */
  uint32_t bin_shift;      /* binning tile of the (first) view: 16 <<
                            * bin_shift pixels (GsplatTuning.bin_auto)  */
  uint32_t depth_sort;     /* how the frame's depth sort ran (the order is
                            * the same): 0 look-back radix passes, 1
                            * reduce-then-scan radix passes, 2 bucket sort
                            * (one view of at most 2^19 slots), 3 none (a
                            * kept order reused)                         */
} GsplatStageTimes;
int32_t gsplat_context_get_stage_times(GsplatContext *ctx, GsplatStageTimes *out);
int32_t gsplat_context_set_stage_timing(GsplatContext *ctx, uint32_t on);

/* ------------------------------------------------------- geometry path */

/*
 * gsplat_surface_renderer_set_geometry_path (gsplat.h:149-157,
 * ffi-c/src/lib.rs:1040-1087) for a context: the geometry path of its scene
 * (as gsplat_scene_create_with_geometry_path). A loaded scene is uploaded
 * again in the new path before the old one is released; on an error the
 * context keeps its path and scene. Default GSPLAT_GEOMETRY_PATH_DIRECT.
 */
int32_t gsplat_context_set_geometry_path(GsplatContext *ctx, uint32_t path);
int32_t gsplat_context_get_geometry_path(const GsplatContext *ctx, uint32_t *out_path);

/* ------------------------------------------------- session order reuse */

/*
 * Opt-in session semantics for a context, after the reference's
 * SurfaceRenderSession (render-wgpu/src/surface_session.rs:17-32, 556-596,
 * 734-744, 851-860, 903-1273). Until gsplat_context_set_sort_interval is
 * called, every frame sorts (the plain SortedAlpha path). After it, each
 * changed camera (gsplat_context_set_camera / _set_auto_camera) is a new
 * camera revision; the depth order is refreshed every `interval` camera
 * changes, and other frames draw the current camera's projection in the last
 * refreshed order, as long as that order lags at most 2 revisions and its
 * camera is within the pose envelope of the current one (translation at most
 * 2% of the scene's bounding-box diagonal, rotation at most 0.01 rad);
 * otherwise the frame refreshes (flag bit 7). Such frames are judged by the
 * reference's stale-order image gate (mean |delta RGB| <= 1/255, at most
 * 0.1% of pixels off by more than 3/255; render-wgpu/src/lib.rs:3491-3545).
 * interval 0 -> INVALID_ARGUMENT.
 */
int32_t gsplat_context_set_sort_interval(GsplatContext *ctx, uint32_t interval);
/* The next frame refreshes its order (SurfaceRenderSession::force_sort_refresh). */
int32_t gsplat_context_force_sort_refresh(GsplatContext *ctx);
/* Telemetry of the last frame in GsplatSurfaceSortStats (gsplat.h:69-84):
 * camera_revision, applied_order_revision, presented_order_revision_lag;
 * flags bit 0 order refreshed, bit 1 order uploaded, bit 7 refresh forced by
 * the lag or pose bounds; the async fields stay 0. */
int32_t gsplat_context_get_sort_stats(const GsplatContext *ctx, GsplatSurfaceSortStats *out_stats);

/* ------------------------------------------------ parity/inspection hooks */

/* Depth keys of the last frame for every splat, in splat order:
 * f32 bits of max(depth,0) when near <= depth <= far, else 0
 * (preprocess_visible_into, render-wgpu/src/lib.rs:1263-1296, and the GPU
 * statement direct_gpu_order.wgsl:78-98). `cap` >= splat count. */
int32_t gsplat_context_read_depth_keys(GsplatContext *ctx, uint32_t *out, uint64_t cap);

/* Back-to-front draw order of the last frame: the V visible splat indices
 * sorted by depth key descending, ties by ascending index
 * (CpuSortBackend::sort_values_by_keys, sort/src/lib.rs:36-61). */
int32_t gsplat_context_read_sorted_indices(GsplatContext *ctx, uint32_t *out, uint64_t cap,
                                           uint32_t *out_count);

/* Per-splat projection of the current camera, 12 floats per splat:
 * [drawn, center.x, center.y, axis_u.x, axis_u.y, axis_v.x, axis_v.y,
 *  r*a, g*a, b*a, a, 0] in NDC -- the quantities of project_splat and
 * eval_color (splat_surface_direct.wgsl:131-217, 328-334); drawn = 0 for
 * splats the shader turns into invalid_splat(). cap >= 12*N floats. */
int32_t gsplat_context_read_projection(GsplatContext *ctx, float *out, uint64_t cap);

/* Number of (tile, splat) pairs per 16x16 tile of the last frame, row-major
 * tiles; cap >= tiles_x*tiles_y. */
int32_t gsplat_context_read_tile_counts(GsplatContext *ctx, uint32_t *out, uint64_t cap,
                                        uint32_t *out_tiles_x, uint32_t *out_tiles_y);

/* --------------------------------------------------------- standalone sort */

/* GPU statement of CpuSortBackend::sort_values_by_keys
 * (sort/src/lib.rs:36-61): stable sort of `values` by `keys` descending;
 * equal keys keep their input order. Host buffers; runs on the device. */
int32_t gsplat_sort_values_by_keys(const uint32_t *keys, uint32_t *values, uint64_t len);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc). The same
 * sort as gsplat_sort_values_by_keys, same arguments and result, with every
 * device buffer (keys, values, their radix alternates, the sort's status
 * words) in one fixed 18 MiB device array of the library that each GPU it
 * runs on holds until the process ends: no hipMalloc, hipFree or stream
 * creation. len > 1048576 (2^20) -> UNSUPPORTED. Calls share their GPU's
 * array: the calls of one process are serialised by a lock and have finished
 * on the device when they return; overlapping use of the array from outside
 * these calls would race. */
int32_t gsplat_sort_values_by_keys_noalloc(const uint32_t *keys, uint32_t *values, uint64_t len);

/* ------------------------------------------------------------- PLY loader */

/* Resource budgets of gsplat_io_ply::PlyLoadLimits (io-ply/src/lib.rs:28-47);
 * defaults 1 GiB input, 1 MiB header, 5,000,000 vertices, 128 properties,
 * 1 GiB decoded scene. */
typedef struct GsplatPlyLimits {
  uint64_t max_input_bytes;
  uint64_t max_header_bytes;
  uint64_t max_vertices;
  uint64_t max_vertex_properties;
  uint64_t max_scene_bytes;
} GsplatPlyLimits;

typedef struct GsplatPlyScene GsplatPlyScene;

GsplatPlyLimits gsplat_ply_limits_default(void);
/* load_ply / parse_ply_bytes_with_limits (io-ply/src/lib.rs:178-294); error
 * codes as PlyLoadError::code (:161-176), detail "<fn>: <PlyLoadError text>". */
int32_t gsplat_ply_load(const char *path, GsplatPlyScene **out_scene);
int32_t gsplat_ply_parse_bytes(const uint8_t *bytes, uint64_t len, const GsplatPlyLimits *limits,
                               GsplatPlyScene **out_scene);
/* parse_ply_text_with_limits: additionally rejects non-ASCII formats. */
int32_t gsplat_ply_parse_text(const uint8_t *bytes, uint64_t len, const GsplatPlyLimits *limits,
                              GsplatPlyScene **out_scene);
/* Borrowed view valid until gsplat_ply_free. */
int32_t gsplat_ply_scene_view(const GsplatPlyScene *scene, GsplatSceneView *out_view);
void gsplat_ply_free(GsplatPlyScene *scene);

/* 1 when a gfx950 device is visible, else 0. */
uint32_t gsplat_device_available(void);

/* ------------------------------------------------------- resident scenes */

/*
 * A scene uploaded once to HBM (load-time precompute of Renderer::load_scene,
 * render-wgpu/src/lib.rs:488-497, kept on the device) and shared by any
 * number of batches. Validation and error codes are those of
 * gsplat_context_load_scene_memory / _load_scene_path. The scene may be
 * destroyed while a batch that rendered it still holds its last frame: the
 * device memory is released when no batch references it.
 */
typedef struct GsplatScene GsplatScene;

typedef struct GsplatSceneInfo {
  uint32_t count;
  uint32_t sh_degree;
  uint64_t device_bytes;  /* HBM held by the scene's device image */
  uint32_t geometry_path; /* GsplatGeometryPath the scene was created with */
  uint32_t reserved;
} GsplatSceneInfo;

int32_t gsplat_scene_create(const GsplatSceneView *scene, GsplatScene **out_scene);
int32_t gsplat_scene_create_from_path(const char *path, GsplatScene **out_scene);
/*
 * A resident scene in a chosen geometry path (the reference's
 * gsplat_surface_renderer_create_*_with_geometry_path, gsplat.h:119-140, for
 * a scene). GSPLAT_GEOMETRY_PATH_DIRECT (0) is the path above;
 * GSPLAT_GEOMETRY_PATH_PACKED_ATLAS (1) holds the scene as the reference's
 * packed records (render-wgpu/src/packed_atlas.rs): 20 bytes per splat
 * (u16 bounds-relative position, u8 opacity, three u10 log-scales, a
 * smallest-three quaternion, the RGB10 degree-0 colour) plus, for SH degrees
 * 1..3, 48 bytes of i8 SH coefficients over per-channel scene scales, decoded
 * by the projection. Lossy: its frames match the direct path's within the
 * reference's packed-vs-direct gate (mean |dRGB| <= 1/255, at most 0.1% of
 * pixels off by more than 3/255; render-wgpu/src/lib.rs:3520-3537,
 * 3733-3771). SH degree 4 is GSPLAT_ERROR_UNSUPPORTED in the packed path,
 * and so is GSPLAT_GEOMETRY_PATH_PAGED_ACTIVE_ATLAS (2); other values are
 * GSPLAT_ERROR_INVALID_ARGUMENT.
 */
int32_t gsplat_scene_create_with_geometry_path(const GsplatSceneView *scene, uint32_t geometry_path,
                                               GsplatScene **out_scene);
int32_t gsplat_scene_create_from_path_with_geometry_path(const char *path, uint32_t geometry_path,
                                                         GsplatScene **out_scene);
/*
 * The packed records of GSPLAT_GEOMETRY_PATH_PACKED_ATLAS for a scene, on
 * the host (pack_scene, render-wgpu/src/packed_atlas.rs:420-426, 479-566):
 * word w of splat i at out_words[w * count + i]; w 0..4 the hot record
 * (x | y << 16, z | opacity << 16, three u10 log-scales, smallest-three
 * quaternion, RGB10 colour), w 5..16 the 48-byte SH sidecar (byte
 * 15 * channel + k is coefficient k of the channel as i8; SH degrees 1..3
 * only). out_consts receives GSPLAT_PACKED_CONSTS floats: bounds min xyz,
 * bounds span xyz, log-scale min, log-scale span, SH scales rgb, five zeros,
 * then the coverage cut of each of the 256 opacity codes. *out_words_per_splat
 * = 5 or 17. With out_words NULL only *out_words_per_splat is set; otherwise
 * words_cap must hold words_per_splat * count words and consts_cap
 * GSPLAT_PACKED_CONSTS floats (GSPLAT_ERROR_INVALID_ARGUMENT otherwise).
 * SH degree 4: GSPLAT_ERROR_UNSUPPORTED. Needs no device.
 */
#define GSPLAT_PACKED_CONSTS 272
int32_t gsplat_pack_scene(const GsplatSceneView *scene, uint32_t *out_words, uint64_t words_cap, float *out_consts,
                          uint64_t consts_cap, uint32_t *out_words_per_splat);
int32_t gsplat_scene_get_info(const GsplatScene *scene, GsplatSceneInfo *out_info);
void gsplat_scene_destroy(GsplatScene *scene);

/* ---------------------------------------------------------------- batches */

/*
 * One view of a batch: a resident scene, a camera (validated as in
 * gsplat_context_set_camera) and a target size (as gsplat_context_set_size).
 * Each view's result is the frame a context holding that scene, camera and
 * size would render (Renderer::render_frame, render-wgpu/src/lib.rs:723-742),
 * identical whatever else the batch holds.
 */
typedef struct GsplatViewDesc {
  const GsplatScene *scene;
  GsplatCamera camera;
  uint32_t width;
  uint32_t height;
} GsplatViewDesc;

/* Renders up to 256 views per frame on its own device stream: one projection
 * launch, one segmented depth sort, one tile binning and one raster launch
 * for all views. A batch is confined to one thread at a time, like a
 * context. Scenes in a batch of more than one view hold at most 2^24-1
 * splats each. */
typedef struct GsplatBatch GsplatBatch;

int32_t gsplat_batch_create(GsplatBatch **out_batch);
void gsplat_batch_destroy(GsplatBatch *batch);
/* Enqueues the frame and returns without waiting for the device. */
int32_t gsplat_batch_render(GsplatBatch *batch, const GsplatViewDesc *views, uint32_t count);
int32_t gsplat_batch_wait(GsplatBatch *batch);
/* Per-view results of the last frame (wait implied); view >= count ->
 * INVALID_ARGUMENT, no frame -> SCENE_NOT_LOADED. */
int32_t gsplat_batch_readback_rgba8(GsplatBatch *batch, uint32_t view, uint8_t *out, uint64_t len);
int32_t gsplat_batch_readback_rgba32f(GsplatBatch *batch, uint32_t view, float *out, uint64_t len);
int32_t gsplat_batch_get_view_stats(GsplatBatch *batch, uint32_t view, GsplatStats *out_stats);
int32_t gsplat_batch_read_sorted_indices(GsplatBatch *batch, uint32_t view, uint32_t *out, uint64_t cap,
                                         uint32_t *out_count);
int32_t gsplat_batch_read_tile_counts(GsplatBatch *batch, uint32_t view, uint32_t *out, uint64_t cap,
                                      uint32_t *out_tiles_x, uint32_t *out_tiles_y);
/* Whole-frame stage times: splat_count and visible_count are sums over the
 * views, tiles_x the total tile count over the views and tiles_y 1. */
int32_t gsplat_batch_get_stage_times(GsplatBatch *batch, GsplatStageTimes *out);
int32_t gsplat_batch_set_stage_timing(GsplatBatch *batch, uint32_t on);

/* ----------------------------------------------------------------- tuning */

/*
 * Launch configuration of a context or batch. Every setting renders
 * bit-identical images; they trade work between binning, sorting and
 * blending:
 *   bin_shift           splats are binned into tiles of (16 << bin_shift)
 *                       pixels, 0..3 (the largest, with bin_auto); each
 *                       16x16 raster tile keeps the splats of its binning
 *                       tile that overlap it
 *   raster_variant      0 per-lane exit, 1 wave-uniform loop (default),
 *                       2 wave-uniform loop with two pixels per thread
 *                       (its views bin into tiles of at least 32 pixels),
 *                       3 wave-uniform loop in a persistent grid taking
 *                       the (view, tile) work items from per-XCD queues,
 *                       4 8x8 raster tiles: each wave blends one 8x8 block
 *                       of a 16x16 workgroup over the splats that meet it
 *   sort_rts_min_tiles  radix sorts of at least this many 4096-key tiles use
 *                       reduce-then-scan passes, shorter ones look-back passes
 *   frame_ms            (autotune output) median device time of a frame
 *   blend_min           a raster tile blends its staged splats once this
 *                       many (1..256, default 64) are staged or its list
 *                       ends; smaller stops a saturated tile's list walk
 *                       sooner, larger runs longer blend loops
 *   bin_auto            1 (default): each view bins into tiles of about three
 *                       on-screen splat diameters, at most 16 << bin_shift
 *                       pixels (small views of large scenes get small
 *                       binning tiles); 0: every view uses bin_shift
 *   rows2_min_tiles     with raster_variant 1, frames of at least this many
 *                       16x16 tiles whose views share one binning tile of at
 *                       least 32 pixels, with at most one splat per two
 *                       pixels, blend as variant 2 (default 24576, three
 *                       1080p views; 0: never)
 */
typedef struct GsplatTuning {
  uint32_t bin_shift;
  uint32_t raster_variant;
  uint32_t sort_rts_min_tiles;
  float frame_ms;
  uint32_t blend_min;
  uint32_t bin_auto;
  uint32_t rows2_min_tiles;
} GsplatTuning;

int32_t gsplat_batch_get_tuning(GsplatBatch *batch, GsplatTuning *out_tuning);
int32_t gsplat_batch_set_tuning(GsplatBatch *batch, const GsplatTuning *tuning);
/* Renders `views` `reps` times under each candidate (bin_shift 0..3 x
 * sort_rts_min_tiles {128, 256, 1024} x raster_variant {1, 2, 3, 4}, variant 2
 * from bin_shift 1, then blend_min {32, 128, 256} around the fastest of
 * those), keeps the fastest and returns it;
 * the batch's last frame is then those views under the kept tuning. */
int32_t gsplat_batch_autotune(GsplatBatch *batch, const GsplatViewDesc *views, uint32_t count, uint32_t reps,
                              GsplatTuning *out_best);
int32_t gsplat_context_get_tuning(GsplatContext *ctx, GsplatTuning *out_tuning);
int32_t gsplat_context_set_tuning(GsplatContext *ctx, const GsplatTuning *tuning);
/* As gsplat_batch_autotune for the context's scene, camera and size. */
int32_t gsplat_context_autotune(GsplatContext *ctx, uint32_t reps, GsplatTuning *out_best);

/* ----------------------------------------------------------------- server */

/*
 * Request loop over resident scenes with continuous batching: requests that
 * arrive while batches run on the device are folded into the next batch as
 * soon as one of its pipelines (config.pipelines) is free, and consecutive
 * batches overlap on the device. Each result is the frame a context would
 * render for the request alone. RGBA8 results for targets allocated with
 * gsplat_host_alloc are written there by the raster itself. Admission is by device memory (the analogue of the
 * reference's direct_scene_preflight, render-wgpu/src/lib.rs:2202-2304, with
 * a byte budget in place of the storage-binding limit): resident scenes plus
 * each pipeline's batch working set stay within memory_budget_bytes. All
 * gsplat_server_* calls are thread-safe; destroy must not race with them.
 */
typedef struct GsplatServer GsplatServer;

typedef struct GsplatServerConfig {
  uint64_t memory_budget_bytes; /* 0: 90% of the device memory free at create */
  uint32_t max_batch_views;     /* views folded into one batch, 1..256 */
  uint32_t max_queue;           /* queued requests before submit refuses */
  uint32_t pipelines;           /* batches in flight at once, 1..8 (default 3) */
} GsplatServerConfig;

typedef struct GsplatRenderRequest {
  uint32_t scene_id;
  GsplatCamera camera;
  uint32_t width;
  uint32_t height;
  uint8_t *out_rgba8; /* width*height*4 bytes or NULL; kept allocated and unused
                       * by the caller until gsplat_server_wait returns */
  uint64_t out_len;
} GsplatRenderRequest;

typedef struct GsplatRequestResult {
  int32_t status;         /* GsplatErrorCode of the request */
  uint32_t visible_count; /* V of the view */
  uint32_t batch_views;   /* views in the batch that rendered it */
  float queue_ms;         /* submit -> batch start */
  float device_ms;        /* device time of that batch */
  float total_ms;         /* submit -> image delivered */
} GsplatRequestResult;

typedef struct GsplatServerStats {
  uint64_t submitted;
  uint64_t completed;
  uint64_t rejected;
  uint64_t batches;
  uint64_t views_in_batches;
  uint64_t resident_scene_bytes;
  uint64_t memory_budget_bytes;
  uint32_t scenes;
  uint32_t queued;
  uint64_t overflow_rerenders; /* batches rendered again after their tile-pair
                                * buffers overflowed (the buffers then grow) */
  uint64_t buffer_grows;       /* batches whose pipeline had to reallocate its
                                * frame buffers first (in its stream's order) */
} GsplatServerStats;

/* {0, 64, 4096} */
GsplatServerConfig gsplat_server_config_default(void);
int32_t gsplat_server_create(const GsplatServerConfig *config, GsplatServer **out_server);
/* Waits for batches on the device; requests still queued end with INTERNAL. */
void gsplat_server_destroy(GsplatServer *server);
/* Registers a resident scene (at most 2^24-1 splats); UNSUPPORTED when the
 * budget cannot hold it plus one 1080p view of it per pipeline. */
int32_t gsplat_server_add_scene(GsplatServer *server, const GsplatScene *scene, uint32_t *out_scene_id);
int32_t gsplat_server_remove_scene(GsplatServer *server, uint32_t scene_id);
/* Queues a render. The RGBA8 result goes to out_rgba8 (may be NULL). When
 * out_rgba8 lies in a gsplat_host_alloc block, the raster writes the pixels
 * there directly from the device; other memory receives a copy after the
 * batch. Either way the buffer must stay allocated, and must not be reused,
 * until gsplat_server_wait returns for the ticket (gsplat_host_free refuses a
 * block that a queued or running request still targets). */
int32_t gsplat_server_submit(GsplatServer *server, const GsplatRenderRequest *request, uint64_t *out_ticket);
/* Blocks until the request is done and releases the ticket; returns the
 * request's own status (its message in gsplat_last_error_message). */
int32_t gsplat_server_wait(GsplatServer *server, uint64_t ticket, GsplatRequestResult *out_result);
int32_t gsplat_server_get_stats(GsplatServer *server, GsplatServerStats *out_stats);

/* Pinned host memory for server outputs (NULL on failure). gsplat_host_free
 * returns INVALID_ARGUMENT, freeing nothing, for a pointer that
 * gsplat_host_alloc did not return or a block that a queued or running
 * request still targets; NULL is a no-op. */
void *gsplat_host_alloc(uint64_t bytes);
int32_t gsplat_host_free(void *ptr);

/* Camera at `position` looking at `target` with the reference's look-at
 * basis and quaternion (camera_rotation_looking_at / quat_from_camera_basis,
 * ffi-c/src/lib.rs:1677-1739); fov/near/far are taken as given. */
GsplatCamera gsplat_camera_look_at(const float position[3], const float target[3],
                                   float vertical_fov_radians, float near_plane, float far_plane);

#ifdef __cplusplus
}
#endif

#endif /* GSPLAT_FFI_C_GSPLAT_EXT_H */
