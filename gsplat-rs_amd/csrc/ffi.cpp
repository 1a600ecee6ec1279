// ffi.cpp -- the gsplat.h / gsplat_ext.h C ABI.
//
// Mirrors crates/gsplat-ffi-c/src/lib.rs: thread-local last-error message
// ("ok" on success, "<fn>: <detail>" on failure, :190-223), static code
// strings (:267-283), NULL checks before any work, and an exception firewall
// on every entry point standing in for catch_unwind (:225-255).

#include <cmath>
#include <chrono>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "common.h"
#include "renderer.h"
#include "server.h"
#include "sort.h"
#include "sort_noalloc.h"

struct GsplatContext {
  std::unique_ptr<gs::Renderer> renderer;
  GsplatCamera camera;
};

struct GsplatPlyScene {
  gs::HostScene scene;
};

struct GsplatScene {
  std::shared_ptr<gs::DeviceScene> scene;
};

struct GsplatBatch {
  gs::Pipeline pipe;
  float host_ms = 0.0f;
};

struct GsplatServer {
  std::unique_ptr<gs::Server> server;
};

namespace {

thread_local std::string g_last_error = "ok";

void set_last(const std::string &m) {
  std::string clean;
  clean.reserve(m.size());
  for (char c : m)
    if (c != '\0') clean.push_back(c);
  g_last_error = clean;
}

int32_t ok() {
  set_last("ok");
  return 0;
}

int32_t fail(gs::Code c, const std::string &msg) {
  set_last(msg);
  return (int32_t)c;
}

// catch_unwind equivalent: gs::Error keeps its code, anything else is a
// panic (INTERNAL with the fixed text).
template <typename F>
int32_t guard(const char *op, F &&body) {
  try {
    return body();
  } catch (const gs::Error &e) {
    return fail(e.code, std::string(op) + ": " + e.what());
  } catch (const std::bad_alloc &) {
    return fail(gs::Code::Internal, std::string(op) + ": panic caught at C ABI boundary");
  } catch (...) {
    try {
      set_last(std::string(op) + ": panic caught at C ABI boundary");
    } catch (...) {
    }
    return (int32_t)gs::Code::Internal;
  }
}

int32_t null_arg(const char *op, const char *what) {
  return fail(gs::Code::InvalidArgument, std::string(op) + ": " + what + " is null");
}

gs::PlyLimits limits_from(const GsplatPlyLimits *l) {
  gs::PlyLimits out;
  if (l) {
    out.max_input_bytes = l->max_input_bytes;
    out.max_header_bytes = l->max_header_bytes;
    out.max_vertices = l->max_vertices;
    out.max_vertex_properties = l->max_vertex_properties;
    out.max_scene_bytes = l->max_scene_bytes;
  }
  return out;
}

// CStr::to_str: a path must be UTF-8 (ffi-c/src/lib.rs, load_scene_path).
bool path_is_utf8(const char *path) {
  const size_t len = std::strlen(path);
  const unsigned char *u = reinterpret_cast<const unsigned char *>(path);
  size_t i = 0;
  while (i < len) {
    const unsigned char c = u[i];
    const size_t l = c < 0x80 ? 1 : (c & 0xE0) == 0xC0 ? 2 : (c & 0xF0) == 0xE0 ? 3 : (c & 0xF8) == 0xF0 ? 4 : 0;
    bool good = l != 0 && i + l <= len;
    for (size_t k = 1; good && k < l; ++k) good = (u[i + k] & 0xC0) == 0x80;
    if (good && l == 2) good = c >= 0xC2;
    if (!good) return false;
    i += l;
  }
  return true;
}

}  // namespace

extern "C" {

// ------------------------------------------------------------ gsplat.h

uint32_t gsplat_version_major(void) { return 0; }
uint32_t gsplat_version_minor(void) { return 1; }

const char *gsplat_error_message(int32_t code) {
  switch (code) {
    case 0: return "ok";
    case 1: return "invalid argument";
    case 2: return "not found";
    case 3: return "parse failed";
    case 4: return "unsupported";
    case 5: return "scene not loaded";
    case 100: return "internal error";
    default: return "unknown error";
  }
}

const char *gsplat_last_error_message(void) { return g_last_error.c_str(); }

GsplatConfig gsplat_config_default(void) {
  GsplatConfig c;
  c.width = 1280;
  c.height = 720;
  c.mode = GSPLAT_RENDER_MODE_SORTED_ALPHA;
  return c;
}

GsplatCamera gsplat_camera_default(void) { return gs::camera_default(); }

// gsplat_context_create (ffi-c/src/lib.rs:325-375).
int32_t gsplat_context_create(GsplatConfig config, GsplatContext **out_ctx) {
  static const char *op = "gsplat_context_create";
  return guard(op, [&]() -> int32_t {
    if (!out_ctx) return fail(gs::Code::InvalidArgument, "gsplat_context_create: out_ctx is null");
    *out_ctx = nullptr;
    if (config.mode != GSPLAT_RENDER_MODE_SORTED_ALPHA)
      return fail(gs::Code::InvalidArgument, "gsplat_context_create: unsupported render mode");
    if (config.width == 0 || config.height == 0)
      return fail(gs::Code::InvalidArgument, "gsplat_context_create: invalid renderer configuration");
    auto ctx = std::make_unique<GsplatContext>();
    ctx->renderer = std::make_unique<gs::Renderer>(config.width, config.height);
    ctx->camera = gs::camera_default();
    *out_ctx = ctx.release();
    return ok();
  });
}

void gsplat_context_destroy(GsplatContext *ctx) {
  try {
    delete ctx;
  } catch (...) {
    set_last("gsplat_context_destroy: panic caught at C ABI boundary");
  }
}

int32_t gsplat_context_set_camera(GsplatContext *ctx, GsplatCamera camera) {
  static const char *op = "gsplat_context_set_camera";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    if (!gs::camera_valid(camera)) return fail(gs::Code::InvalidArgument, std::string(op) + ": invalid camera");
    ctx->camera = camera;
    ctx->renderer->camera_set(camera);
    return ok();
  });
}

int32_t gsplat_context_set_auto_camera(GsplatContext *ctx) {
  static const char *op = "gsplat_context_set_auto_camera";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    try {
      ctx->camera = gs::auto_camera(ctx->renderer->scene(), ctx->renderer->width(), ctx->renderer->height());
    } catch (const gs::Error &e) {
      return fail(e.code, std::string(op) + ": scene is not loaded");
    }
    ctx->renderer->camera_set(ctx->camera);
    return ok();
  });
}

int32_t gsplat_context_load_scene_path(GsplatContext *ctx, const char *path) {
  static const char *op = "gsplat_context_load_scene_path";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    if (!path) return null_arg(op, "path");
    if (!path_is_utf8(path)) return fail(gs::Code::InvalidArgument, std::string(op) + ": path is not valid UTF-8");
    gs::HostScene scene = gs::ply_load(path, gs::PlyLimits());
    ctx->renderer->load_scene(std::move(scene));
    return ok();
  });
}

int32_t gsplat_context_render_frame(GsplatContext *ctx) {
  static const char *op = "gsplat_context_render_frame";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    ctx->renderer->render(ctx->camera);
    return ok();
  });
}

// SurfaceRenderSession::set_sort_interval (render-wgpu/src/surface_session.rs:
// 851-860) and the FFI's check (ffi-c/src/lib.rs:1005-1033), for a context.
int32_t gsplat_context_set_sort_interval(GsplatContext *ctx, uint32_t interval) {
  static const char *op = "gsplat_context_set_sort_interval";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    if (interval == 0) return fail(gs::Code::InvalidArgument, std::string(op) + ": interval must be positive");
    ctx->renderer->set_sort_interval(interval);
    return ok();
  });
}

// gsplat_surface_renderer_set_geometry_path (ffi-c/src/lib.rs:1040-1087) for a
// context: 0 direct, 1 packed atlas, 2 (paged active atlas) unsupported, any
// other value INVALID_ARGUMENT "unsupported path".
int32_t gsplat_context_set_geometry_path(GsplatContext *ctx, uint32_t path) {
  static const char *op = "gsplat_context_set_geometry_path";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    if (path > 2) return fail(gs::Code::InvalidArgument, std::string(op) + ": unsupported path");
    ctx->renderer->set_geometry_path(path);
    return ok();
  });
}

int32_t gsplat_context_get_geometry_path(const GsplatContext *ctx, uint32_t *out_path) {
  static const char *op = "gsplat_context_get_geometry_path";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    if (!out_path) return null_arg(op, "out_path");
    *out_path = ctx->renderer->geometry_path();
    return ok();
  });
}

int32_t gsplat_context_force_sort_refresh(GsplatContext *ctx) {
  static const char *op = "gsplat_context_force_sort_refresh";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    ctx->renderer->force_sort_refresh();
    return ok();
  });
}

// gsplat_surface_renderer_get_sort_stats (ffi-c/src/lib.rs:1566-1591) for a context.
int32_t gsplat_context_get_sort_stats(const GsplatContext *ctx, GsplatSurfaceSortStats *out_stats) {
  static const char *op = "gsplat_context_get_sort_stats";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    if (!out_stats) return null_arg(op, "out_stats");
    *out_stats = ctx->renderer->sort_stats();
    return ok();
  });
}

int32_t gsplat_context_get_stats(const GsplatContext *ctx, GsplatStats *out_stats) {
  static const char *op = "gsplat_context_get_stats";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    if (!out_stats) return null_arg(op, "out_stats");
    *out_stats = ctx->renderer->stats();
    return ok();
  });
}

// ------------------------------------------- surface API (headless device)

#define GS_SURFACE_UNSUPPORTED(op) \
  fail(gs::Code::Unsupported, std::string(op) + ": native window surfaces are not available on this device")

int32_t gsplat_surface_renderer_create_android(void *native_window, const char *path, uint32_t, uint32_t,
                                               GsplatSurfaceRenderer **out_renderer) {
  static const char *op = "gsplat_surface_renderer_create_android";
  return guard(op, [&]() -> int32_t {
    if (!out_renderer) return null_arg(op, "out_renderer");
    *out_renderer = nullptr;
    if (!native_window) return null_arg(op, "native_window");
    if (!path) return null_arg(op, "path");
    return GS_SURFACE_UNSUPPORTED(op);
  });
}

int32_t gsplat_surface_renderer_create_android_with_geometry_path(void *native_window, const char *path, uint32_t w,
                                                                  uint32_t h, uint32_t,
                                                                  GsplatSurfaceRenderer **out_renderer) {
  return gsplat_surface_renderer_create_android(native_window, path, w, h, out_renderer);
}

int32_t gsplat_surface_renderer_create_uikit(void *ui_view, void *, const char *path, uint32_t, uint32_t,
                                             GsplatSurfaceRenderer **out_renderer) {
  static const char *op = "gsplat_surface_renderer_create_uikit";
  return guard(op, [&]() -> int32_t {
    if (!out_renderer) return null_arg(op, "out_renderer");
    *out_renderer = nullptr;
    if (!ui_view) return null_arg(op, "ui_view");
    if (!path) return null_arg(op, "path");
    return GS_SURFACE_UNSUPPORTED(op);
  });
}

int32_t gsplat_surface_renderer_create_uikit_with_geometry_path(void *ui_view, void *ui_view_controller,
                                                                const char *path, uint32_t w, uint32_t h, uint32_t,
                                                                GsplatSurfaceRenderer **out_renderer) {
  return gsplat_surface_renderer_create_uikit(ui_view, ui_view_controller, path, w, h, out_renderer);
}

void gsplat_surface_renderer_destroy(GsplatSurfaceRenderer *) {}

#define GS_SURFACE_ENTRY(name, ...)                                  \
  int32_t name(__VA_ARGS__) {                                        \
    static const char *op = #name;                                   \
    return guard(op, [&]() -> int32_t {                              \
      if (!renderer) return null_arg(op, "renderer");                \
      return GS_SURFACE_UNSUPPORTED(op);                             \
    });                                                              \
  }

GS_SURFACE_ENTRY(gsplat_surface_renderer_resize, GsplatSurfaceRenderer *renderer, uint32_t, uint32_t)
GS_SURFACE_ENTRY(gsplat_surface_renderer_set_sort_interval, GsplatSurfaceRenderer *renderer, uint32_t)
GS_SURFACE_ENTRY(gsplat_surface_renderer_set_geometry_path, GsplatSurfaceRenderer *renderer, uint32_t)
GS_SURFACE_ENTRY(gsplat_surface_renderer_set_gpu_preproject, GsplatSurfaceRenderer *renderer, uint32_t)
GS_SURFACE_ENTRY(gsplat_surface_renderer_set_gpu_preproject_double_buffer, GsplatSurfaceRenderer *renderer, uint32_t)
GS_SURFACE_ENTRY(gsplat_surface_renderer_set_async_sort, GsplatSurfaceRenderer *renderer, uint32_t)
GS_SURFACE_ENTRY(gsplat_surface_renderer_set_async_geometry, GsplatSurfaceRenderer *renderer, uint32_t)
GS_SURFACE_ENTRY(gsplat_surface_renderer_set_instance_buffer_count, GsplatSurfaceRenderer *renderer, uint32_t)
GS_SURFACE_ENTRY(gsplat_surface_renderer_set_frame_latency, GsplatSurfaceRenderer *renderer, uint32_t)
GS_SURFACE_ENTRY(gsplat_surface_renderer_reset_camera, GsplatSurfaceRenderer *renderer)
GS_SURFACE_ENTRY(gsplat_surface_renderer_orbit, GsplatSurfaceRenderer *renderer, float, float)
GS_SURFACE_ENTRY(gsplat_surface_renderer_zoom, GsplatSurfaceRenderer *renderer, float)
GS_SURFACE_ENTRY(gsplat_surface_renderer_pan, GsplatSurfaceRenderer *renderer, float, float)
GS_SURFACE_ENTRY(gsplat_surface_renderer_render_frame, GsplatSurfaceRenderer *renderer)

int32_t gsplat_surface_renderer_get_stats(const GsplatSurfaceRenderer *renderer, GsplatStats *out_stats) {
  static const char *op = "gsplat_surface_renderer_get_stats";
  return guard(op, [&]() -> int32_t {
    if (!renderer) return null_arg(op, "renderer");
    if (!out_stats) return null_arg(op, "out_stats");
    return GS_SURFACE_UNSUPPORTED(op);
  });
}

int32_t gsplat_surface_renderer_get_sort_stats(const GsplatSurfaceRenderer *renderer,
                                               GsplatSurfaceSortStats *out_stats) {
  static const char *op = "gsplat_surface_renderer_get_sort_stats";
  return guard(op, [&]() -> int32_t {
    if (!renderer) return null_arg(op, "renderer");
    if (!out_stats) return null_arg(op, "out_stats");
    return GS_SURFACE_UNSUPPORTED(op);
  });
}

// --------------------------------------------------------- gsplat_ext.h

int32_t gsplat_context_load_scene_memory(GsplatContext *ctx, const GsplatSceneView *view) {
  static const char *op = "gsplat_context_load_scene_memory";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    if (!view) return null_arg(op, "scene");
    gs::HostScene s;
    try {
      s = gs::scene_from_view(*view);
    } catch (const gs::Error &e) {
      // Renderer::load_scene maps invalid buffers to InvalidScene
      // (render-wgpu/src/lib.rs:488-490 -> InvalidArgument), except the
      // unsupported SH degree.
      if (e.code == gs::Code::Unsupported) throw;
      return fail(gs::Code::InvalidArgument, std::string(op) + ": invalid scene buffers");
    }
    ctx->renderer->load_scene(std::move(s));
    return ok();
  });
}

int32_t gsplat_context_set_size(GsplatContext *ctx, uint32_t width, uint32_t height) {
  static const char *op = "gsplat_context_set_size";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    ctx->renderer->set_size(width, height);
    return ok();
  });
}

int32_t gsplat_context_get_camera(const GsplatContext *ctx, GsplatCamera *out_camera) {
  static const char *op = "gsplat_context_get_camera";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    if (!out_camera) return null_arg(op, "out_camera");
    *out_camera = ctx->camera;
    return ok();
  });
}

int32_t gsplat_context_wait(GsplatContext *ctx) {
  static const char *op = "gsplat_context_wait";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    ctx->renderer->wait();
    return ok();
  });
}

int32_t gsplat_context_readback_rgba8(GsplatContext *ctx, uint8_t *out, uint64_t len) {
  static const char *op = "gsplat_context_readback_rgba8";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    if (!out) return null_arg(op, "out");
    ctx->renderer->readback_rgba8(out, len);
    return ok();
  });
}

int32_t gsplat_context_readback_rgba32f(GsplatContext *ctx, float *out, uint64_t len) {
  static const char *op = "gsplat_context_readback_rgba32f";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    if (!out) return null_arg(op, "out");
    ctx->renderer->readback_rgba32f(out, len);
    return ok();
  });
}

int32_t gsplat_context_set_stage_timing(GsplatContext *ctx, uint32_t on) {
  static const char *op = "gsplat_context_set_stage_timing";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    ctx->renderer->set_stage_timing(on != 0);
    return ok();
  });
}

int32_t gsplat_context_get_stage_times(GsplatContext *ctx, GsplatStageTimes *out) {
  static const char *op = "gsplat_context_get_stage_times";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    if (!out) return null_arg(op, "out");
    *out = ctx->renderer->stage_times();
    return ok();
  });
}

int32_t gsplat_context_read_depth_keys(GsplatContext *ctx, uint32_t *out, uint64_t cap) {
  static const char *op = "gsplat_context_read_depth_keys";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    if (!out) return null_arg(op, "out");
    ctx->renderer->read_depth_keys(out, cap);
    return ok();
  });
}

int32_t gsplat_context_read_sorted_indices(GsplatContext *ctx, uint32_t *out, uint64_t cap, uint32_t *out_count) {
  static const char *op = "gsplat_context_read_sorted_indices";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    if (!out) return null_arg(op, "out");
    if (!out_count) return null_arg(op, "out_count");
    ctx->renderer->read_sorted(out, cap, out_count);
    return ok();
  });
}

int32_t gsplat_context_read_projection(GsplatContext *ctx, float *out, uint64_t cap) {
  static const char *op = "gsplat_context_read_projection";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    if (!out) return null_arg(op, "out");
    ctx->renderer->read_projection(ctx->camera, out, cap);
    return ok();
  });
}

int32_t gsplat_context_read_tile_counts(GsplatContext *ctx, uint32_t *out, uint64_t cap, uint32_t *out_tiles_x,
                                        uint32_t *out_tiles_y) {
  static const char *op = "gsplat_context_read_tile_counts";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    if (!out) return null_arg(op, "out");
    if (!out_tiles_x) return null_arg(op, "out_tiles_x");
    if (!out_tiles_y) return null_arg(op, "out_tiles_y");
    ctx->renderer->read_tile_counts(out, cap, out_tiles_x, out_tiles_y);
    return ok();
  });
}

// CpuSortBackend::sort_values_by_keys (sort/src/lib.rs:36-61) on the device:
// length mismatch cannot occur through this signature; len <= 1 is a no-op.
int32_t gsplat_sort_values_by_keys(const uint32_t *keys, uint32_t *values, uint64_t len) {
  static const char *op = "gsplat_sort_values_by_keys";
  return guard(op, [&]() -> int32_t {
    if (len <= 1) return ok();
    if (!keys) return null_arg(op, "keys");
    if (!values) return null_arg(op, "values");
    if (len > 0xFFFFFFF0ull) return fail(gs::Code::Unsupported, std::string(op) + ": length exceeds 32-bit indexing");
    if (!gs::device_present()) return fail(gs::Code::Unsupported, std::string(op) + ": gpu backend unavailable");
    const uint32_t n = (uint32_t)len;
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
      return fail(gs::Code::Internal, std::string(op) + ": stream creation failed");
    struct StreamGuard {
      hipStream_t s;
      ~StreamGuard() {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
      }
    } sg{st};
    gs::DevBuf<uint32_t> k, v, ka, va, scratch;
    k.ensure(n);
    v.ensure(n);
    ka.ensure(n);
    va.ensure(n);
    scratch.ensure(gs::onesweep_status_words(n, 32));  // the sort clears what it needs on `st`
    if (hipMemcpyAsync(k.get(), keys, len * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(v.get(), values, len * 4, hipMemcpyHostToDevice, st) != hipSuccess)
      return fail(gs::Code::Internal, std::string(op) + ": upload failed");
    uint32_t *pk = k.get(), *pv = v.get(), *pka = ka.get(), *pva = va.get();
    gs::onesweep_sort(&pk, &pv, &pka, &pva, n, nullptr, 32, true, scratch.get(), scratch.capacity(), st);
    if (hipMemcpyAsync(values, pv, len * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail(gs::Code::Internal, std::string(op) + ": download failed");
    return ok();
  });
}

// The same sort with every device buffer in a fixed device array of the
// library (no allocation); lengths above 2^20 are refused.
int32_t gsplat_sort_values_by_keys_noalloc(const uint32_t *keys, uint32_t *values, uint64_t len) {
  static const char *op = "gsplat_sort_values_by_keys_noalloc";
  return guard(op, [&]() -> int32_t {
    if (len <= 1) return ok();
    if (!keys) return null_arg(op, "keys");
    if (!values) return null_arg(op, "values");
    if (len > gs::kNoAllocSortMaxLen)
      return fail(gs::Code::Unsupported, std::string(op) + ": length " + std::to_string(len) +
                                             " exceeds the no-allocation limit of " +
                                             std::to_string(gs::kNoAllocSortMaxLen));
    if (!gs::device_present()) return fail(gs::Code::Unsupported, std::string(op) + ": gpu backend unavailable");
    gs::sort_values_by_keys_noalloc(keys, values, (uint32_t)len);
    return ok();
  });
}

GsplatPlyLimits gsplat_ply_limits_default(void) {
  gs::PlyLimits d;
  GsplatPlyLimits l;
  l.max_input_bytes = d.max_input_bytes;
  l.max_header_bytes = d.max_header_bytes;
  l.max_vertices = d.max_vertices;
  l.max_vertex_properties = d.max_vertex_properties;
  l.max_scene_bytes = d.max_scene_bytes;
  return l;
}

int32_t gsplat_ply_load(const char *path, GsplatPlyScene **out_scene) {
  static const char *op = "gsplat_ply_load";
  return guard(op, [&]() -> int32_t {
    if (!out_scene) return null_arg(op, "out_scene");
    *out_scene = nullptr;
    if (!path) return null_arg(op, "path");
    auto s = std::make_unique<GsplatPlyScene>();
    s->scene = gs::ply_load(path, gs::PlyLimits());
    *out_scene = s.release();
    return ok();
  });
}

static int32_t ply_parse_entry(const char *op, const uint8_t *bytes, uint64_t len, const GsplatPlyLimits *limits,
                               GsplatPlyScene **out_scene, bool text) {
  return guard(op, [&]() -> int32_t {
    if (!out_scene) return null_arg(op, "out_scene");
    *out_scene = nullptr;
    if (!bytes && len) return null_arg(op, "bytes");
    static const uint8_t empty = 0;
    auto s = std::make_unique<GsplatPlyScene>();
    s->scene = gs::ply_parse(bytes ? bytes : &empty, (size_t)len, limits_from(limits), text);
    *out_scene = s.release();
    return ok();
  });
}

int32_t gsplat_ply_parse_bytes(const uint8_t *bytes, uint64_t len, const GsplatPlyLimits *limits,
                               GsplatPlyScene **out_scene) {
  return ply_parse_entry("gsplat_ply_parse_bytes", bytes, len, limits, out_scene, false);
}

int32_t gsplat_ply_parse_text(const uint8_t *bytes, uint64_t len, const GsplatPlyLimits *limits,
                              GsplatPlyScene **out_scene) {
  return ply_parse_entry("gsplat_ply_parse_text", bytes, len, limits, out_scene, true);
}

int32_t gsplat_ply_scene_view(const GsplatPlyScene *scene, GsplatSceneView *out_view) {
  static const char *op = "gsplat_ply_scene_view";
  return guard(op, [&]() -> int32_t {
    if (!scene) return null_arg(op, "scene");
    if (!out_view) return null_arg(op, "out_view");
    const gs::HostScene &s = scene->scene;
    out_view->count = s.count;
    out_view->sh_degree = s.sh_degree;
    out_view->positions = s.positions.data();
    out_view->opacity = s.opacity.data();
    out_view->scale_xyz = s.scale_xyz.data();
    out_view->rotation_xyzw = s.rotation_xyzw.data();
    out_view->color_dc = s.color_dc.data();
    out_view->sh_rest = s.has_sh_rest ? s.sh_rest.data() : nullptr;
    return ok();
  });
}

void gsplat_ply_free(GsplatPlyScene *scene) { delete scene; }

GsplatCamera gsplat_camera_look_at(const float position[3], const float target[3], float fov, float near_plane,
                                   float far_plane) {
  GsplatCamera c = gs::camera_default();
  if (!position || !target) return c;
  for (int a = 0; a < 3; ++a) c.position[a] = position[a];
  gs::rotation_looking_at(position, target, c.rotation_xyzw);
  c.vertical_fov_radians = fov;
  c.near_plane = near_plane;
  c.far_plane = far_plane;
  return c;
}

uint32_t gsplat_device_available(void) {
  try {
    return gs::device_present() ? 1u : 0u;
  } catch (...) {
    return 0u;
  }
}

// ------------------------------------------------------- resident scenes

}  // extern "C"

namespace {

int32_t make_scene(gs::HostScene &&s, GsplatScene **out, uint32_t geometry_path = 0) {
  if (!gs::device_present()) throw gs::Error(gs::Code::Unsupported, "gpu rasterizer unavailable");
  gs::validate_scene(s);
  auto h = std::make_unique<GsplatScene>();
  h->scene = gs::upload_scene(s, nullptr, geometry_path);
  *out = h.release();
  return ok();
}

int32_t create_scene(const char *op, const GsplatSceneView *view, uint32_t geometry_path, GsplatScene **out_scene) {
  return guard(op, [&]() -> int32_t {
    if (!out_scene) return null_arg(op, "out_scene");
    *out_scene = nullptr;
    if (!view) return null_arg(op, "scene");
    if (geometry_path > 2) return fail(gs::Code::InvalidArgument, std::string(op) + ": unsupported path");
    gs::HostScene s;
    try {
      s = gs::scene_from_view(*view);
    } catch (const gs::Error &e) {
      if (e.code == gs::Code::Unsupported) throw;
      return fail(gs::Code::InvalidArgument, std::string(op) + ": invalid scene buffers");
    }
    return make_scene(std::move(s), out_scene, geometry_path);
  });
}

int32_t create_scene_from_path(const char *op, const char *path, uint32_t geometry_path, GsplatScene **out_scene) {
  return guard(op, [&]() -> int32_t {
    if (!out_scene) return null_arg(op, "out_scene");
    *out_scene = nullptr;
    if (!path) return null_arg(op, "path");
    if (geometry_path > 2) return fail(gs::Code::InvalidArgument, std::string(op) + ": unsupported path");
    if (!path_is_utf8(path)) return fail(gs::Code::InvalidArgument, std::string(op) + ": path is not valid UTF-8");
    return make_scene(gs::ply_load(path, gs::PlyLimits()), out_scene, geometry_path);
  });
}

}  // namespace

extern "C" {

int32_t gsplat_scene_create(const GsplatSceneView *view, GsplatScene **out_scene) {
  return create_scene("gsplat_scene_create", view, 0, out_scene);
}

int32_t gsplat_scene_create_from_path(const char *path, GsplatScene **out_scene) {
  return create_scene_from_path("gsplat_scene_create_from_path", path, 0, out_scene);
}

int32_t gsplat_pack_scene(const GsplatSceneView *view, uint32_t *out_words, uint64_t words_cap, float *out_consts,
                          uint64_t consts_cap, uint32_t *out_words_per_splat) {
  static const char *op = "gsplat_pack_scene";
  static_assert(GSPLAT_PACKED_CONSTS == gs::kPackedConsts, "packed constant count");
  return guard(op, [&]() -> int32_t {
    if (!view) return null_arg(op, "scene");
    if (!out_words_per_splat) return null_arg(op, "out_words_per_splat");
    gs::HostScene s;
    try {
      s = gs::scene_from_view(*view);
    } catch (const gs::Error &e) {
      if (e.code == gs::Code::Unsupported) throw;
      return fail(gs::Code::InvalidArgument, std::string(op) + ": invalid scene buffers");
    }
    const uint32_t degree = s.has_sh_rest ? s.sh_degree : 0;
    if (degree > 3) return fail(gs::Code::Unsupported, std::string(op) + ": the packed scene codec holds SH degrees 0..3");
    *out_words_per_splat = gs::packed_words(degree);
    if (!out_words) return ok();
    if (!out_consts) return null_arg(op, "out_consts");
    if (words_cap < (uint64_t)gs::packed_words(degree) * s.count || consts_cap < gs::kPackedConsts)
      return fail(gs::Code::InvalidArgument, std::string(op) + ": output buffers too small");
    const gs::PackedScene p = gs::pack_scene(s);
    std::memcpy(out_words, p.words.data(), p.words.size() * sizeof(uint32_t));
    std::memcpy(out_consts, p.consts.data(), p.consts.size() * sizeof(float));
    return ok();
  });
}

int32_t gsplat_scene_create_with_geometry_path(const GsplatSceneView *view, uint32_t geometry_path,
                                               GsplatScene **out_scene) {
  return create_scene("gsplat_scene_create_with_geometry_path", view, geometry_path, out_scene);
}

int32_t gsplat_scene_create_from_path_with_geometry_path(const char *path, uint32_t geometry_path,
                                                         GsplatScene **out_scene) {
  return create_scene_from_path("gsplat_scene_create_from_path_with_geometry_path", path, geometry_path, out_scene);
}

int32_t gsplat_scene_get_info(const GsplatScene *scene, GsplatSceneInfo *out_info) {
  static const char *op = "gsplat_scene_get_info";
  return guard(op, [&]() -> int32_t {
    if (!scene) return null_arg(op, "scene");
    if (!out_info) return null_arg(op, "out_info");
    out_info->count = scene->scene->count;
    out_info->sh_degree = scene->scene->sh_degree;
    out_info->device_bytes = scene->scene->bytes();
    out_info->geometry_path = scene->scene->geometry_path;
    out_info->reserved = 0;
    return ok();
  });
}

void gsplat_scene_destroy(GsplatScene *scene) {
  try {
    delete scene;
  } catch (...) {
  }
}

// ---------------------------------------------------------------- batches

}  // extern "C"

namespace {

// The ABI's fields over `base` (settings the ABI does not carry stay).
gs::Tuning tuning_from(const GsplatTuning &t, gs::Tuning base) {
  gs::Tuning out = base;
  out.bin_shift = t.bin_shift;
  if (t.raster_variant > 4) throw gs::Error(gs::Code::InvalidArgument, "raster_variant must be 0 to 4");
  out.raster = static_cast<gs::RasterVariant>(t.raster_variant);
  out.rts_min_tiles = t.sort_rts_min_tiles;
  out.blend_min = t.blend_min;
  out.bin_auto = t.bin_auto != 0 ? 1u : 0u;
  out.rows2_min_tiles = t.rows2_min_tiles;
  return out;
}

GsplatTuning tuning_to(const gs::Tuning &t, float ms) {
  GsplatTuning out;
  out.bin_shift = t.bin_shift;
  out.raster_variant = static_cast<uint32_t>(t.raster);
  out.sort_rts_min_tiles = t.rts_min_tiles;
  out.frame_ms = ms;
  out.blend_min = t.blend_min;
  out.bin_auto = t.bin_auto != 0 ? 1u : 0u;
  out.rows2_min_tiles = t.rows2_min_tiles;
  return out;
}

std::vector<gs::ViewSpec> view_specs(const char *op, const GsplatViewDesc *views, uint32_t count) {
  if (count == 0 || count > gs::kMaxViews)
    throw gs::Error(gs::Code::InvalidArgument, "view count must be 1.." + std::to_string(gs::kMaxViews));
  std::vector<gs::ViewSpec> specs;
  specs.reserve(count);
  for (uint32_t v = 0; v < count; ++v) {
    const GsplatViewDesc &d = views[v];
    const std::string at = "view " + std::to_string(v) + ": ";
    if (!d.scene) throw gs::Error(gs::Code::InvalidArgument, at + "scene is null");
    if (!gs::camera_valid(d.camera)) throw gs::Error(gs::Code::InvalidArgument, at + "invalid camera");
    if (d.width == 0 || d.height == 0) throw gs::Error(gs::Code::InvalidArgument, at + "invalid renderer configuration");
    if (d.width > 16384 || d.height > 16384)
      throw gs::Error(gs::Code::Unsupported, at + "render dimensions exceed the device 2D texture limit 16384");
    specs.push_back(gs::ViewSpec{d.scene->scene,
                                 gs::make_frame_params(d.camera.position, d.camera.rotation_xyzw,
                                                       d.camera.vertical_fov_radians, d.camera.near_plane,
                                                       d.camera.far_plane, d.width, d.height)});
  }
  (void)op;
  return specs;
}

}  // namespace

extern "C" {


int32_t gsplat_batch_create(GsplatBatch **out_batch) {
  static const char *op = "gsplat_batch_create";
  return guard(op, [&]() -> int32_t {
    if (!out_batch) return null_arg(op, "out_batch");
    *out_batch = nullptr;
    auto b = std::make_unique<GsplatBatch>();
    b->pipe.set_raster_variant(gs::raster_variant_from_env());
    *out_batch = b.release();
    return ok();
  });
}

void gsplat_batch_destroy(GsplatBatch *batch) {
  try {
    delete batch;
  } catch (...) {
  }
}

int32_t gsplat_batch_render(GsplatBatch *batch, const GsplatViewDesc *views, uint32_t count) {
  static const char *op = "gsplat_batch_render";
  return guard(op, [&]() -> int32_t {
    if (!batch) return null_arg(op, "batch");
    if (!views) return null_arg(op, "views");
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<gs::ViewSpec> specs = view_specs(op, views, count);
    batch->pipe.enqueue(std::move(specs));
    batch->host_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return ok();
  });
}

int32_t gsplat_batch_wait(GsplatBatch *batch) {
  static const char *op = "gsplat_batch_wait";
  return guard(op, [&]() -> int32_t {
    if (!batch) return null_arg(op, "batch");
    batch->pipe.wait();
    return ok();
  });
}

int32_t gsplat_batch_readback_rgba8(GsplatBatch *batch, uint32_t view, uint8_t *out, uint64_t len) {
  static const char *op = "gsplat_batch_readback_rgba8";
  return guard(op, [&]() -> int32_t {
    if (!batch) return null_arg(op, "batch");
    if (!out) return null_arg(op, "out");
    batch->pipe.readback_rgba8(view, out, len);
    return ok();
  });
}

int32_t gsplat_batch_readback_rgba32f(GsplatBatch *batch, uint32_t view, float *out, uint64_t len) {
  static const char *op = "gsplat_batch_readback_rgba32f";
  return guard(op, [&]() -> int32_t {
    if (!batch) return null_arg(op, "batch");
    if (!out) return null_arg(op, "out");
    batch->pipe.readback_rgba32f(view, out, len);
    return ok();
  });
}

int32_t gsplat_batch_get_view_stats(GsplatBatch *batch, uint32_t view, GsplatStats *out_stats) {
  static const char *op = "gsplat_batch_get_view_stats";
  return guard(op, [&]() -> int32_t {
    if (!batch) return null_arg(op, "batch");
    if (!out_stats) return null_arg(op, "out_stats");
    const uint32_t vis = batch->pipe.visible(view);
    const gs::FrameTimes t = batch->pipe.times();
    out_stats->frame_ms = batch->host_ms;
    out_stats->preprocess_ms = t.project_ms;
    out_stats->sort_ms = t.sort_ms;
    out_stats->raster_ms = t.bin_ms + t.raster_ms;
    out_stats->visible_count = vis;
    out_stats->drawn_count = vis;
    return ok();
  });
}

int32_t gsplat_batch_read_sorted_indices(GsplatBatch *batch, uint32_t view, uint32_t *out, uint64_t cap,
                                         uint32_t *out_count) {
  static const char *op = "gsplat_batch_read_sorted_indices";
  return guard(op, [&]() -> int32_t {
    if (!batch) return null_arg(op, "batch");
    if (!out_count) return null_arg(op, "out_count");
    if (!out && cap) return null_arg(op, "out");
    batch->pipe.read_sorted(view, out, cap, out_count);
    return ok();
  });
}

int32_t gsplat_batch_read_tile_counts(GsplatBatch *batch, uint32_t view, uint32_t *out, uint64_t cap,
                                      uint32_t *out_tiles_x, uint32_t *out_tiles_y) {
  static const char *op = "gsplat_batch_read_tile_counts";
  return guard(op, [&]() -> int32_t {
    if (!batch) return null_arg(op, "batch");
    if (!out) return null_arg(op, "out");
    if (!out_tiles_x || !out_tiles_y) return null_arg(op, "out_tiles");
    batch->pipe.read_tile_counts(view, out, cap, out_tiles_x, out_tiles_y);
    return ok();
  });
}

int32_t gsplat_batch_set_stage_timing(GsplatBatch *batch, uint32_t on) {
  static const char *op = "gsplat_batch_set_stage_timing";
  return guard(op, [&]() -> int32_t {
    if (!batch) return null_arg(op, "batch");
    batch->pipe.set_stage_timing(on != 0);
    return ok();
  });
}

int32_t gsplat_batch_get_stage_times(GsplatBatch *batch, GsplatStageTimes *out) {
  static const char *op = "gsplat_batch_get_stage_times";
  return guard(op, [&]() -> int32_t {
    if (!batch) return null_arg(op, "batch");
    if (!out) return null_arg(op, "out");
    *out = GsplatStageTimes{};
    if (!batch->pipe.has_frame()) return ok();
    const gs::FrameTimes t = batch->pipe.times();
    out->project_ms = t.project_ms;
    out->sort_ms = t.sort_ms;
    out->bin_ms = t.bin_ms;
    out->raster_ms = t.raster_ms;
    out->total_ms = t.total_ms;
    out->splat_count = (uint32_t)std::min<uint64_t>(batch->pipe.total_splats(), 0xFFFFFFFFu);
    uint64_t vis = 0;
    for (uint32_t v = 0; v < batch->pipe.view_count(); ++v) vis += batch->pipe.visible(v);
    out->visible_count = (uint32_t)std::min<uint64_t>(vis, 0xFFFFFFFFu);
    out->tiles_x = batch->pipe.total_tiles();
    out->tiles_y = 1;
    out->tile_pairs = batch->pipe.pairs();
    out->blend_evals = t.blend_evals;
    out->blend_tests = t.blend_tests;
    out->bin_pairs = batch->pipe.bin_pairs();
    out->bin_shift = batch->pipe.view(0).bin_shift;
    out->depth_sort = batch->pipe.depth_sort_mode();
    return ok();
  });
}

// ----------------------------------------------------------------- tuning


int32_t gsplat_batch_get_tuning(GsplatBatch *batch, GsplatTuning *out_tuning) {
  static const char *op = "gsplat_batch_get_tuning";
  return guard(op, [&]() -> int32_t {
    if (!batch) return null_arg(op, "batch");
    if (!out_tuning) return null_arg(op, "out_tuning");
    *out_tuning = tuning_to(batch->pipe.tuning(), 0.0f);
    return ok();
  });
}

int32_t gsplat_batch_set_tuning(GsplatBatch *batch, const GsplatTuning *tuning) {
  static const char *op = "gsplat_batch_set_tuning";
  return guard(op, [&]() -> int32_t {
    if (!batch) return null_arg(op, "batch");
    if (!tuning) return null_arg(op, "tuning");
    batch->pipe.set_tuning(tuning_from(*tuning, batch->pipe.tuning()));
    return ok();
  });
}

int32_t gsplat_batch_autotune(GsplatBatch *batch, const GsplatViewDesc *views, uint32_t count, uint32_t reps,
                              GsplatTuning *out_best) {
  static const char *op = "gsplat_batch_autotune";
  return guard(op, [&]() -> int32_t {
    if (!batch) return null_arg(op, "batch");
    if (!views) return null_arg(op, "views");
    if (!out_best) return null_arg(op, "out_best");
    const std::vector<gs::TuneSample> samples = batch->pipe.autotune(view_specs(op, views, count), reps);
    gs::TuneSample best = samples.front();
    for (const gs::TuneSample &s : samples)
      if (s.frame_ms < best.frame_ms) best = s;
    *out_best = tuning_to(best.tuning, best.frame_ms);
    return ok();
  });
}

int32_t gsplat_context_get_tuning(GsplatContext *ctx, GsplatTuning *out_tuning) {
  static const char *op = "gsplat_context_get_tuning";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    if (!out_tuning) return null_arg(op, "out_tuning");
    *out_tuning = tuning_to(ctx->renderer->tuning(), 0.0f);
    return ok();
  });
}

int32_t gsplat_context_set_tuning(GsplatContext *ctx, const GsplatTuning *tuning) {
  static const char *op = "gsplat_context_set_tuning";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    if (!tuning) return null_arg(op, "tuning");
    ctx->renderer->set_tuning(tuning_from(*tuning, ctx->renderer->tuning()));
    return ok();
  });
}

int32_t gsplat_context_autotune(GsplatContext *ctx, uint32_t reps, GsplatTuning *out_best) {
  static const char *op = "gsplat_context_autotune";
  return guard(op, [&]() -> int32_t {
    if (!ctx) return null_arg(op, "ctx");
    if (!out_best) return null_arg(op, "out_best");
    const gs::TuneSample best = ctx->renderer->autotune(ctx->camera, reps);
    *out_best = tuning_to(best.tuning, best.frame_ms);
    return ok();
  });
}

// ----------------------------------------------------------------- server

GsplatServerConfig gsplat_server_config_default(void) {
  GsplatServerConfig c;
  c.memory_budget_bytes = 0;
  c.max_batch_views = 64;
  c.max_queue = 4096;
  c.pipelines = 3;
  return c;
}

int32_t gsplat_server_create(const GsplatServerConfig *config, GsplatServer **out_server) {
  static const char *op = "gsplat_server_create";
  return guard(op, [&]() -> int32_t {
    if (!out_server) return null_arg(op, "out_server");
    *out_server = nullptr;
    if (!config) return null_arg(op, "config");
    gs::ServerConfig c;
    c.memory_budget = config->memory_budget_bytes;
    c.max_batch_views = config->max_batch_views;
    c.max_queue = config->max_queue;
    c.pipelines = config->pipelines;
    auto s = std::make_unique<GsplatServer>();
    s->server = std::make_unique<gs::Server>(c);
    *out_server = s.release();
    return ok();
  });
}

void gsplat_server_destroy(GsplatServer *server) {
  try {
    delete server;
  } catch (...) {
  }
}

int32_t gsplat_server_add_scene(GsplatServer *server, const GsplatScene *scene, uint32_t *out_scene_id) {
  static const char *op = "gsplat_server_add_scene";
  return guard(op, [&]() -> int32_t {
    if (!server) return null_arg(op, "server");
    if (!scene) return null_arg(op, "scene");
    if (!out_scene_id) return null_arg(op, "out_scene_id");
    *out_scene_id = server->server->add_scene(scene->scene);
    return ok();
  });
}

int32_t gsplat_server_remove_scene(GsplatServer *server, uint32_t scene_id) {
  static const char *op = "gsplat_server_remove_scene";
  return guard(op, [&]() -> int32_t {
    if (!server) return null_arg(op, "server");
    server->server->remove_scene(scene_id);
    return ok();
  });
}

int32_t gsplat_server_submit(GsplatServer *server, const GsplatRenderRequest *request, uint64_t *out_ticket) {
  static const char *op = "gsplat_server_submit";
  return guard(op, [&]() -> int32_t {
    if (!server) return null_arg(op, "server");
    if (!request) return null_arg(op, "request");
    if (!out_ticket) return null_arg(op, "out_ticket");
    *out_ticket = server->server->submit(request->scene_id, request->camera, request->width, request->height,
                                         request->out_rgba8, request->out_len);
    return ok();
  });
}

int32_t gsplat_server_wait(GsplatServer *server, uint64_t ticket, GsplatRequestResult *out_result) {
  static const char *op = "gsplat_server_wait";
  return guard(op, [&]() -> int32_t {
    if (!server) return null_arg(op, "server");
    if (!out_result) return null_arg(op, "out_result");
    const gs::RequestResult r = server->server->wait(ticket);
    out_result->status = r.status;
    out_result->visible_count = r.visible;
    out_result->batch_views = r.batch_views;
    out_result->queue_ms = r.queue_ms;
    out_result->device_ms = r.device_ms;
    out_result->total_ms = r.total_ms;
    if (r.status != 0) return fail((gs::Code)r.status, std::string(op) + ": " + r.error);
    return ok();
  });
}

int32_t gsplat_server_get_stats(GsplatServer *server, GsplatServerStats *out_stats) {
  static const char *op = "gsplat_server_get_stats";
  return guard(op, [&]() -> int32_t {
    if (!server) return null_arg(op, "server");
    if (!out_stats) return null_arg(op, "out_stats");
    const gs::ServerStats s = server->server->stats();
    out_stats->submitted = s.submitted;
    out_stats->completed = s.completed;
    out_stats->rejected = s.rejected;
    out_stats->batches = s.batches;
    out_stats->views_in_batches = s.views_in_batches;
    out_stats->resident_scene_bytes = s.resident_scene_bytes;
    out_stats->memory_budget_bytes = s.memory_budget;
    out_stats->scenes = s.scenes;
    out_stats->queued = s.queued;
    out_stats->overflow_rerenders = s.overflow_rerenders;
    out_stats->buffer_grows = s.buffer_grows;
    return ok();
  });
}

void *gsplat_host_alloc(uint64_t bytes) {
  try {
    return gs::host_alloc(bytes);
  } catch (const gs::Error &e) {
    set_last(std::string("gsplat_host_alloc: ") + e.what());
    return nullptr;
  } catch (...) {
    set_last("gsplat_host_alloc: panic caught at C ABI boundary");
    return nullptr;
  }
}

int32_t gsplat_host_free(void *ptr) {
  static const char *op = "gsplat_host_free";
  return guard(op, [&]() -> int32_t {
    gs::host_free(ptr);
    return ok();
  });
}

}  // extern "C"
