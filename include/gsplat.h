/*
 * gsplat.h -- C ABI of the MI355X-native SortedAlpha renderer.
 *
 * Drop-in for the gsplat-rs v0.1 C ABI: every symbol, enum value and
 * by-value struct layout below matches crates/gsplat-ffi-c/include/gsplat.h
 * of the reference (symbols :97-201, structs :53-92, codes :28-36), so a
 * caller compiled against that header links and runs against
 * libgsplat_ffi_c.so built from this repository.
 *
 * Semantics kept from the reference:
 *  - every int32_t return is a GsplatErrorCode;
 *  - handles are opaque, owned by the caller after a successful create and
 *    released by the matching destroy, which accepts NULL;
 *  - a handle is confined to one thread at a time (the caller serialises);
 *  - gsplat_last_error_message() is thread-local, "ok" after a success and
 *    "<function>: <detail>" after a failure;
 *  - C++ exceptions never cross the boundary: they become
 *    GSPLAT_ERROR_INTERNAL with "<function>: panic caught at C ABI boundary".
 *
 * Rendering runs on one AMD Instinct MI355X (gfx950) through hand-written
 * HIP kernels. The gsplat_surface_renderer_* entry points need a native
 * window system and return GSPLAT_ERROR_UNSUPPORTED on this headless device
 * (after the same NULL-handle checks as the reference).
 *
 * Additive, MI355X-specific entry points (in-memory scenes, readback,
 * stage timings, the standalone depth sort, batched and served rendering)
 * are declared in gsplat_ext.h.
 */
#ifndef GSPLAT_FFI_C_GSPLAT_H
#define GSPLAT_FFI_C_GSPLAT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSPLAT_API_VERSION_MAJOR_VALUE 0
#define GSPLAT_API_VERSION_MINOR_VALUE 1

/* Reference: gsplat.h:28-36, core/src/lib.rs:154-164. */
typedef enum GsplatErrorCode {
  GSPLAT_OK = 0,
  GSPLAT_ERROR_INVALID_ARGUMENT = 1,
  GSPLAT_ERROR_NOT_FOUND = 2,
  GSPLAT_ERROR_PARSE_FAILED = 3,
  GSPLAT_ERROR_UNSUPPORTED = 4,
  GSPLAT_ERROR_SCENE_NOT_LOADED = 5,
  GSPLAT_ERROR_INTERNAL = 100,
} GsplatErrorCode;

/* Reference: gsplat.h:38-41. SortedAlpha is the only accepted mode. */
typedef enum GsplatRenderMode {
  GSPLAT_RENDER_MODE_SORTED_ALPHA = 0,
} GsplatRenderMode;

/* Reference: gsplat.h:43-51 (surface geometry knob; surface API only). */
typedef enum GsplatGeometryPath {
  GSPLAT_GEOMETRY_PATH_DIRECT = 0,
  GSPLAT_GEOMETRY_PATH_PACKED_ATLAS = 1,
  GSPLAT_GEOMETRY_PATH_PAGED_ACTIVE_ATLAS = 2,
} GsplatGeometryPath;

/* Reference: gsplat.h:53-58. Default 1280x720, mode 0. */
typedef struct GsplatConfig {
  uint32_t width;
  uint32_t height;
  uint32_t mode;
} GsplatConfig;

/* Reference: gsplat.h:60-67, core/src/lib.rs:172-199.
 * preprocess_ms / sort_ms / raster_ms are device times of the projection,
 * depth-sort and (binning + raster) stages of the last frame; frame_ms is
 * the host wall time of gsplat_context_render_frame. */
typedef struct GsplatStats {
  float frame_ms;
  float preprocess_ms;
  float sort_ms;
  float raster_ms;
  uint32_t visible_count;
  uint32_t drawn_count;
} GsplatStats;

/* Reference: gsplat.h:69-84 (surface telemetry; surface API only). */
typedef struct GsplatSurfaceSortStats {
  uint64_t camera_revision;
  uint64_t applied_order_revision;
  uint64_t scheduled_revision;
  uint64_t completed_revision;
  uint32_t presented_order_revision_lag;
  uint32_t observed_result_revision_lag;
  uint32_t flags;
} GsplatSurfaceSortStats;

/* Reference: gsplat.h:86-92. Default: origin, identity (0,0,0,1), 60 deg,
 * near 0.01, far 1000 (core/src/lib.rs:30-72). */
typedef struct GsplatCamera {
  float position[3];
  float rotation_xyzw[4];
  float vertical_fov_radians;
  float near_plane;
  float far_plane;
} GsplatCamera;

typedef struct GsplatContext GsplatContext;
typedef struct GsplatSurfaceRenderer GsplatSurfaceRenderer;

/* Reference: gsplat.h:97-111 (context API; the hot path). */
uint32_t gsplat_version_major(void);
uint32_t gsplat_version_minor(void);
const char *gsplat_error_message(int32_t code);
const char *gsplat_last_error_message(void);

GsplatConfig gsplat_config_default(void);
GsplatCamera gsplat_camera_default(void);

int32_t gsplat_context_create(GsplatConfig config, GsplatContext **out_ctx);
void gsplat_context_destroy(GsplatContext *ctx);
int32_t gsplat_context_set_camera(GsplatContext *ctx, GsplatCamera camera);
int32_t gsplat_context_set_auto_camera(GsplatContext *ctx);
int32_t gsplat_context_load_scene_path(GsplatContext *ctx, const char *path);
int32_t gsplat_context_render_frame(GsplatContext *ctx);
int32_t gsplat_context_get_stats(const GsplatContext *ctx, GsplatStats *out_stats);

/* Reference: gsplat.h:113-201 (window-system surface API).
 * All return GSPLAT_ERROR_UNSUPPORTED here; NULL handles/outputs still
 * return GSPLAT_ERROR_INVALID_ARGUMENT first, as in ffi-c/src/lib.rs. */
int32_t gsplat_surface_renderer_create_android(void *native_window, const char *path,
                                               uint32_t width, uint32_t height,
                                               GsplatSurfaceRenderer **out_renderer);
int32_t gsplat_surface_renderer_create_android_with_geometry_path(
    void *native_window, const char *path, uint32_t width, uint32_t height,
    uint32_t geometry_path, GsplatSurfaceRenderer **out_renderer);
int32_t gsplat_surface_renderer_create_uikit(void *ui_view, void *ui_view_controller,
                                             const char *path, uint32_t width,
                                             uint32_t height,
                                             GsplatSurfaceRenderer **out_renderer);
int32_t gsplat_surface_renderer_create_uikit_with_geometry_path(
    void *ui_view, void *ui_view_controller, const char *path, uint32_t width,
    uint32_t height, uint32_t geometry_path, GsplatSurfaceRenderer **out_renderer);
void gsplat_surface_renderer_destroy(GsplatSurfaceRenderer *renderer);
int32_t gsplat_surface_renderer_resize(GsplatSurfaceRenderer *renderer, uint32_t width,
                                       uint32_t height);
int32_t gsplat_surface_renderer_set_sort_interval(GsplatSurfaceRenderer *renderer,
                                                  uint32_t interval);
int32_t gsplat_surface_renderer_set_geometry_path(GsplatSurfaceRenderer *renderer,
                                                  uint32_t path);
int32_t gsplat_surface_renderer_set_gpu_preproject(GsplatSurfaceRenderer *renderer,
                                                   uint32_t enabled);
int32_t gsplat_surface_renderer_set_gpu_preproject_double_buffer(
    GsplatSurfaceRenderer *renderer, uint32_t enabled);
int32_t gsplat_surface_renderer_set_async_sort(GsplatSurfaceRenderer *renderer,
                                               uint32_t enabled);
int32_t gsplat_surface_renderer_set_async_geometry(GsplatSurfaceRenderer *renderer,
                                                   uint32_t enabled);
int32_t gsplat_surface_renderer_set_instance_buffer_count(GsplatSurfaceRenderer *renderer,
                                                          uint32_t count);
int32_t gsplat_surface_renderer_set_frame_latency(GsplatSurfaceRenderer *renderer,
                                                  uint32_t latency);
int32_t gsplat_surface_renderer_reset_camera(GsplatSurfaceRenderer *renderer);
int32_t gsplat_surface_renderer_orbit(GsplatSurfaceRenderer *renderer,
                                      float delta_yaw_radians, float delta_pitch_radians);
int32_t gsplat_surface_renderer_zoom(GsplatSurfaceRenderer *renderer, float distance_scale);
int32_t gsplat_surface_renderer_pan(GsplatSurfaceRenderer *renderer,
                                    float normalized_delta_x, float normalized_delta_y);
int32_t gsplat_surface_renderer_render_frame(GsplatSurfaceRenderer *renderer);
int32_t gsplat_surface_renderer_get_stats(const GsplatSurfaceRenderer *renderer,
                                          GsplatStats *out_stats);
int32_t gsplat_surface_renderer_get_sort_stats(const GsplatSurfaceRenderer *renderer,
                                               GsplatSurfaceSortStats *out_stats);

#ifdef __cplusplus
}
#endif

#endif /* GSPLAT_FFI_C_GSPLAT_H */
