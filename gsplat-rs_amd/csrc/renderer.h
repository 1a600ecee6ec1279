// renderer.h -- one-GPU SortedAlpha renderer behind a GsplatContext.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <memory>
#include <vector>

#include "common.h"
#include "kernels.h"

namespace gs {

// Grow-only device allocation.
template <typename T>
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void ensure(size_t count);
  // As ensure(), clearing the allocation to zero whenever it is (re)made.
  void ensure_zeroed(size_t count);
  void release();
  T *get() const { return ptr_; }
  size_t capacity() const { return cap_; }

 private:
  T *ptr_ = nullptr;
  size_t cap_ = 0;
};

// Scene resident in HBM in the coefficient-major SoA layout of
// prepare_scene(); shared by renderers, batches and the server.
struct DeviceScene {
  uint32_t count = 0;
  uint32_t sh_degree = 0;
  DevBuf<float> soa;
  uint64_t bytes() const { return (uint64_t)soa_fields(sh_degree) * count * sizeof(float); }
};
std::shared_ptr<DeviceScene> upload_scene(const HostScene &s, hipStream_t stream);

// Per-view working set of one frame: depth keys, sort buffers, records,
// tile pairs, tile ranges and the RGBA targets.
struct ViewWork {
  DevBuf<uint32_t> keys, vals, keys_alt, vals_alt;
  DevBuf<float4> rec;
  DevBuf<uint2> rect;
  DevBuf<uint32_t> counts, offsets;
  DevBuf<uint32_t> pair_keys, pair_vals, pair_keys_alt, pair_vals_alt;
  DevBuf<uint2> ranges;
  DevBuf<float4> img32;
  DevBuf<uchar4> img8;
  DevBuf<uint32_t> sort_scratch, partials;
  DevBuf<uint64_t> total;
  DevBuf<uint32_t> visible, m_dev, info;
  uint32_t pair_cap = 0;
  uint32_t *sorted_vals = nullptr;  // back-to-front splat order of the last frame
  uint32_t *tile_list = nullptr;    // tile-sorted splat ids of the last frame
};

class Renderer {
 public:
  Renderer(uint32_t width, uint32_t height);
  ~Renderer();

  void set_size(uint32_t width, uint32_t height);
  uint32_t width() const { return width_; }
  uint32_t height() const { return height_; }
  void load_scene(HostScene &&scene);
  const HostScene *scene() const { return host_scene_.get(); }

  // Enqueues a full frame for `camera` and returns without waiting for the
  // device (as the reference's queue.submit, render-wgpu/src/lib.rs:2751).
  void render(const GsplatCamera &camera);
  // Waits for the last frame; re-renders it when the tile pairs overflowed
  // their buffers (which are then grown).
  void wait();
  GsplatStats stats();
  GsplatStageTimes stage_times();
  uint64_t overflow_rerenders() const { return overflow_rerenders_; }

  void readback_rgba8(uint8_t *out, uint64_t len);
  void readback_rgba32f(float *out, uint64_t len);
  void read_depth_keys(uint32_t *out, uint64_t cap);
  void read_sorted(uint32_t *out, uint64_t cap, uint32_t *count);
  void read_projection(const GsplatCamera &camera, float *out, uint64_t cap);
  void read_tile_counts(uint32_t *out, uint64_t cap, uint32_t *tx, uint32_t *ty);

 private:
  void enqueue_frame(const GsplatCamera &camera);
  bool absorb_frame_info();  // true when the frame's pairs overflowed
  void ensure_frame_buffers(uint32_t n);
  void ensure_target();
  void ensure_pairs(uint32_t capacity, uint32_t tile_bits);

  hipStream_t stream_ = nullptr;
  uint32_t width_, height_;
  std::unique_ptr<HostScene> host_scene_;
  std::shared_ptr<DeviceScene> scene_;
  ViewWork w_;
  hipEvent_t ev_[6] = {};  // stage boundaries 0..4, 5 = frame summary copied
  uint32_t *pinned_ = nullptr;  // frame summary: [V, M lo, M hi, overflow]
  bool frame_valid_ = false;
  bool frame_pending_ = false;
  GsplatCamera last_cam_{};
  uint32_t frame_w_ = 0, frame_h_ = 0;
  FrameParams last_fp_{};
  uint32_t visible_ = 0;
  uint64_t pairs_ = 0;
  uint64_t overflow_rerenders_ = 0;
  float host_ms_ = 0.0f;
  GsplatStats last_stats_{};
  GsplatStageTimes last_times_{};
  RasterVariant raster_variant_ = RasterVariant::TileLds;
};

bool device_present();

}  // namespace gs
