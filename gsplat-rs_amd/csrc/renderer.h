// renderer.h -- one-GPU SortedAlpha renderer behind a GsplatContext: a
// pipeline rendering a batch of one view.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>

#include "common.h"
#include "order_reuse.h"
#include "pipeline.h"

namespace gs {

class Renderer {
 public:
  Renderer(uint32_t width, uint32_t height);

  void set_size(uint32_t width, uint32_t height);
  uint32_t width() const { return width_; }
  uint32_t height() const { return height_; }
  void load_scene(HostScene &&scene);
  const HostScene *scene() const { return host_scene_.get(); }
  // Geometry path of the scene on the device (GsplatGeometryPath: 0 direct,
  // 1 packed atlas). A loaded scene is uploaded again in the new path before
  // the old one is released; a failure leaves path and scene unchanged.
  void set_geometry_path(uint32_t path);
  uint32_t geometry_path() const { return geometry_path_; }

  // Enqueues a full frame for `camera` and returns without waiting for the
  // device (as the reference's queue.submit, render-wgpu/src/lib.rs:2751).
  void render(const GsplatCamera &camera);
  // Waits for the last frame (re-rendered when its tile pairs overflowed).
  void wait();
  GsplatStats stats();
  GsplatStageTimes stage_times();
  uint64_t overflow_rerenders() const { return pipe_.overflow_rerenders(); }
  void set_tuning(const Tuning &t) { pipe_.set_tuning(t); }
  void set_stage_timing(bool on) { pipe_.set_stage_timing(on); }
  const Tuning &tuning() const { return pipe_.tuning(); }
  // Autotunes on the scene and `camera` at the current size; the context's
  // last frame is then that frame.
  TuneSample autotune(const GsplatCamera &camera, uint32_t reps);

  void readback_rgba8(uint8_t *out, uint64_t len);
  void readback_rgba32f(float *out, uint64_t len);
  void read_depth_keys(uint32_t *out, uint64_t cap);
  void read_sorted(uint32_t *out, uint64_t cap, uint32_t *count);
  void read_projection(const GsplatCamera &camera, float *out, uint64_t cap);
  void read_tile_counts(uint32_t *out, uint64_t cap, uint32_t *tx, uint32_t *ty);

  // Session order reuse (off until set_sort_interval is called; the plain
  // path sorts every frame). With it, the context refreshes its depth order
  // every `interval` camera changes, within the lag and pose bounds of
  // OrderReusePolicy, and presents the other frames in the kept order.
  void set_sort_interval(uint32_t interval);
  void force_sort_refresh() { policy_.force_refresh(); }
  // The context's camera was set (counts camera revisions).
  void camera_set(const GsplatCamera &c) { policy_.camera_set(c); }
  GsplatSurfaceSortStats sort_stats() const { return policy_.stats(); }

 private:
  void require_frame() const;

  Pipeline pipe_;
  uint32_t width_, height_;
  std::unique_ptr<HostScene> host_scene_;
  std::shared_ptr<DeviceScene> scene_;
  bool frame_valid_ = false;
  float host_ms_ = 0.0f;
  OrderReusePolicy policy_;
  bool reuse_enabled_ = false;
  uint32_t geometry_path_ = 0;
};

// Raster variant chosen by GSPLAT_RASTER_VARIANT: 0 per-lane exit, else the
// default wave-uniform loop unrolled by four.
RasterVariant raster_variant_from_env();

}  // namespace gs
