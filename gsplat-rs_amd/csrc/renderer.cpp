// renderer.cpp -- GsplatContext renderer: one view through the pipeline.

#include "renderer.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <string>

namespace gs {

constexpr uint32_t kMaxDimension = 16384;

RasterVariant raster_variant_from_env() {
  if (const char *v = std::getenv("GSPLAT_RASTER_VARIANT")) {
    const int k = std::atoi(v);
    return k == 0   ? RasterVariant::Linear
           : k == 2 ? RasterVariant::Rows2
           : k == 3 ? RasterVariant::Queue
           : k == 4 ? RasterVariant::Blocks8
                    : RasterVariant::Unrolled;
  }
  return RasterVariant::Unrolled;
}

Renderer::Renderer(uint32_t width, uint32_t height) : width_(width), height_(height) {
  pipe_.set_raster_variant(raster_variant_from_env());
  // A context reports its phase times with every frame, as FrameStats does
  // (render-wgpu/src/lib.rs:565-583): stage events on unless the
  // environment says otherwise (gsplat_context_set_stage_timing turns them off).
  if (!std::getenv("GSPLAT_STAGE_TIMES")) pipe_.set_stage_timing(true);
  set_size(width, height);
}

void Renderer::set_size(uint32_t width, uint32_t height) {
  if (width == 0 || height == 0) throw Error(Code::InvalidArgument, "invalid renderer configuration");
  if (width > kMaxDimension || height > kMaxDimension)
    throw Error(Code::Unsupported, "render dimensions " + std::to_string(width) + "x" + std::to_string(height) +
                                       " exceed the device 2D texture limit " + std::to_string(kMaxDimension));
  width_ = width;
  height_ = height;
}

void Renderer::load_scene(HostScene &&scene) {
  validate_scene(scene);
  pipe_.wait();
  scene_ = upload_scene(scene, pipe_.stream(), geometry_path_);
  // The pose envelope's translation bound: 2% of the bounding-box diagonal
  // (SurfaceRenderSession::new, surface_session.rs:659-674).
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = 0; i < scene.count; ++i)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], scene.positions[3 * (size_t)i + a]);
      hi[a] = std::max(hi[a], scene.positions[3 * (size_t)i + a]);
    }
  float d2 = 0.0f;
  for (int a = 0; a < 3; ++a) d2 += scene.count ? (hi[a] - lo[a]) * (hi[a] - lo[a]) : 0.0f;
  policy_.set_scene_diagonal(std::sqrt(d2));
  host_scene_ = std::make_unique<HostScene>(std::move(scene));
  frame_valid_ = false;
}

void Renderer::set_geometry_path(uint32_t path) {
  if (path > 2) throw Error(Code::InvalidArgument, "unsupported path");
  if (path == 2) throw Error(Code::Unsupported, "the paged active atlas is not built");
  if (path == geometry_path_) return;
  if (host_scene_) {
    pipe_.wait();
    auto next = upload_scene(*host_scene_, pipe_.stream(), path);
    scene_ = std::move(next);
    policy_.force_refresh();
    frame_valid_ = false;
  }
  geometry_path_ = path;
}

void Renderer::set_sort_interval(uint32_t interval) {
  if (interval == 0) throw Error(Code::InvalidArgument, "interval must be positive");
  policy_.set_interval(interval);
  if (!reuse_enabled_) policy_.force_refresh();
  reuse_enabled_ = true;
  pipe_.set_keep_order(true);
}

void Renderer::render(const GsplatCamera &cam) {
  const auto t0 = std::chrono::steady_clock::now();
  if (!scene_) throw Error(Code::SceneNotLoaded, "scene not loaded");
  if (!camera_valid(cam)) throw Error(Code::InvalidArgument, "invalid camera");
  const FrameParams fp = make_frame_params(cam.position, cam.rotation_xyzw, cam.vertical_fov_radians,
                                           cam.near_plane, cam.far_plane, width_, height_);
  if (reuse_enabled_) {
    policy_.camera_set(cam);
    const OrderReusePolicy::Plan p = policy_.plan(pipe_.has_order(scene_.get()));
    pipe_.enqueue({ViewSpec{scene_, fp}}, {}, !p.refresh);
    policy_.finish(p);
  } else {
    pipe_.enqueue({ViewSpec{scene_, fp}});
  }
  frame_valid_ = true;
  host_ms_ = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void Renderer::wait() { pipe_.wait(); }

TuneSample Renderer::autotune(const GsplatCamera &cam, uint32_t reps) {
  if (!scene_) throw Error(Code::SceneNotLoaded, "scene not loaded");
  if (!camera_valid(cam)) throw Error(Code::InvalidArgument, "invalid camera");
  const FrameParams fp = make_frame_params(cam.position, cam.rotation_xyzw, cam.vertical_fov_radians,
                                           cam.near_plane, cam.far_plane, width_, height_);
  const std::vector<TuneSample> samples = pipe_.autotune({ViewSpec{scene_, fp}}, reps);
  frame_valid_ = true;
  TuneSample best = samples.front();
  for (const TuneSample &s : samples)
    if (s.frame_ms < best.frame_ms) best = s;
  return best;
}

void Renderer::require_frame() const {
  if (!frame_valid_) throw Error(Code::SceneNotLoaded, "no rendered frame");
}

GsplatStats Renderer::stats() {
  GsplatStats s{};
  if (!frame_valid_) return s;
  const FrameTimes t = pipe_.times();
  const uint32_t vis = pipe_.visible(0);
  // FrameStats (core/src/lib.rs:172-180): visible = drawn = V on this path
  // (render-wgpu/src/lib.rs:565-583, 736).
  s.frame_ms = host_ms_;
  s.preprocess_ms = t.project_ms;
  s.sort_ms = t.sort_ms;
  s.raster_ms = t.bin_ms + t.raster_ms;
  s.visible_count = vis;
  s.drawn_count = vis;
  return s;
}

GsplatStageTimes Renderer::stage_times() {
  GsplatStageTimes s{};
  if (!frame_valid_) return s;
  const FrameTimes t = pipe_.times();
  const ViewDev &d = pipe_.view(0);
  s.project_ms = t.project_ms;
  s.sort_ms = t.sort_ms;
  s.bin_ms = t.bin_ms;
  s.raster_ms = t.raster_ms;
  s.total_ms = t.total_ms;
  s.splat_count = d.n;
  s.visible_count = pipe_.visible(0);
  s.tiles_x = d.fp.tiles_x;
  s.tiles_y = d.fp.tiles_y;
  s.tile_pairs = pipe_.pairs();
  s.blend_evals = t.blend_evals;
  s.blend_tests = t.blend_tests;
  s.bin_pairs = pipe_.bin_pairs();
  s.bin_shift = d.bin_shift;
  s.depth_sort = pipe_.depth_sort_mode();
  return s;
}

void Renderer::readback_rgba8(uint8_t *out, uint64_t len) {
  if (!frame_valid_) throw Error(Code::SceneNotLoaded, "no rendered frame to read back");
  pipe_.readback_rgba8(0, out, len);
}

void Renderer::readback_rgba32f(float *out, uint64_t len) {
  if (!frame_valid_) throw Error(Code::SceneNotLoaded, "no rendered frame to read back");
  pipe_.readback_rgba32f(0, out, len);
}

void Renderer::read_depth_keys(uint32_t *out, uint64_t cap) {
  require_frame();
  const uint32_t n = scene_->count;
  if (cap < n) throw Error(Code::InvalidArgument, "output buffer smaller than the splat count");
  // The depth sort permutes the key buffers in place, so the keys in splat
  // order are regenerated by the same projection for the frame's camera.
  pipe_.wait();
  pipe_.export_projection(*scene_, pipe_.view(0).fp, out, nullptr);
}

void Renderer::read_sorted(uint32_t *out, uint64_t cap, uint32_t *count) {
  require_frame();
  pipe_.read_sorted(0, out, cap, count);
}

void Renderer::read_projection(const GsplatCamera &cam, float *out, uint64_t cap) {
  if (!scene_) throw Error(Code::SceneNotLoaded, "scene not loaded");
  if (!camera_valid(cam)) throw Error(Code::InvalidArgument, "invalid camera");
  if (cap < 12ull * scene_->count) throw Error(Code::InvalidArgument, "output buffer smaller than 12 floats per splat");
  const FrameParams fp = make_frame_params(cam.position, cam.rotation_xyzw, cam.vertical_fov_radians,
                                           cam.near_plane, cam.far_plane, width_, height_);
  pipe_.export_projection(*scene_, fp, nullptr, out);
}

void Renderer::read_tile_counts(uint32_t *out, uint64_t cap, uint32_t *tx, uint32_t *ty) {
  require_frame();
  pipe_.read_tile_counts(0, out, cap, tx, ty);
}

}  // namespace gs
