// renderer.cpp -- frame pipeline of the one-GPU SortedAlpha renderer.
//
// Per frame (Renderer::render_frame, render-wgpu/src/lib.rs:723-742):
//   project   depth key + cull + projection + SH colour, one thread per splat
//   sort      stable descending radix sort of (key, index) -> back-to-front
//   bin       per-splat tile counts front to back, scan, (tile, splat) pair
//             emission, stable radix sort by tile id, tile ranges
//   raster    16x16-tile front-to-back blend into RGBA32F + RGBA8
// No host synchronisation inside a frame: every size after the projection
// (V, M) is read on the device. The (tile, splat) pair buffers have a host-
// chosen capacity; a frame whose pair count exceeds it drops the excess,
// flags itself, and is re-rendered by wait() after the buffers grow.

#include "renderer.h"

#include "sort.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>

namespace gs {

#define GS_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw Error(Code::Internal, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr uint32_t kMaxDimension = 16384;

template <typename T>
void DevBuf<T>::ensure(size_t count) {
  if (count <= cap_) return;
  release();
  size_t want = std::max<size_t>(count, 1);
  hipError_t e = hipMalloc(reinterpret_cast<void **>(&ptr_), want * sizeof(T));
  if (e != hipSuccess) {
    ptr_ = nullptr;
    throw Error(Code::Internal, std::string("device allocation of ") + std::to_string(want * sizeof(T)) +
                                    " bytes failed: " + hipGetErrorString(e));
  }
  cap_ = want;
}

template <typename T>
void DevBuf<T>::ensure_zeroed(size_t count) {
  if (count <= cap_) return;
  ensure(count);
  GS_HIP(hipMemset(ptr_, 0, cap_ * sizeof(T)));
}

template <typename T>
void DevBuf<T>::release() {
  // Work queued on any stream may still use the old allocation.
  if (ptr_) (void)hipDeviceSynchronize();
  if (ptr_) (void)hipFree(ptr_);
  ptr_ = nullptr;
  cap_ = 0;
}

template class DevBuf<float>;
template class DevBuf<uint32_t>;
template class DevBuf<uint64_t>;
template class DevBuf<float4>;
template class DevBuf<uint2>;
template class DevBuf<uchar4>;

bool device_present() {
  int n = 0;
  return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

std::shared_ptr<DeviceScene> upload_scene(const HostScene &s, hipStream_t stream) {
  PreparedScene p = prepare_scene(s);
  auto d = std::make_shared<DeviceScene>();
  d->count = p.count;
  d->sh_degree = p.sh_degree;
  d->soa.ensure(p.soa.size());
  if (!p.soa.empty()) {
    GS_HIP(hipMemcpyAsync(d->soa.get(), p.soa.data(), p.soa.size() * sizeof(float), hipMemcpyHostToDevice, stream));
    GS_HIP(hipStreamSynchronize(stream));
  }
  return d;
}

Renderer::Renderer(uint32_t width, uint32_t height) : width_(width), height_(height) {
  if (!device_present()) throw Error(Code::Unsupported, "gpu rasterizer unavailable");
  GS_HIP(hipSetDevice(0));
  GS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  for (auto &e : ev_) GS_HIP(hipEventCreate(&e));
  GS_HIP(hipHostMalloc(reinterpret_cast<void **>(&pinned_), 4 * sizeof(uint32_t), hipHostMallocDefault));
  w_.visible.ensure(1);
  w_.total.ensure(1);
  w_.m_dev.ensure(1);
  w_.info.ensure(4);
  if (const char *v = std::getenv("GSPLAT_RASTER_VARIANT")) raster_variant_ = (RasterVariant)(std::atoi(v) % 3);
  set_size(width, height);
}

Renderer::~Renderer() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto &e : ev_)
    if (e) (void)hipEventDestroy(e);
  if (pinned_) (void)hipHostFree(pinned_);
  if (stream_) (void)hipStreamDestroy(stream_);
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
  wait();
  scene_ = upload_scene(scene, stream_);
  host_scene_ = std::make_unique<HostScene>(std::move(scene));
  frame_valid_ = false;
  w_.pair_cap = 0;
}

void Renderer::ensure_frame_buffers(uint32_t n) {
  const size_t m = std::max<uint32_t>(n, 1);
  w_.keys.ensure(m);
  w_.vals.ensure(m);
  w_.keys_alt.ensure(m);
  w_.vals_alt.ensure(m);
  w_.rec.ensure(kRecordFloat4 * m);
  w_.rect.ensure(m);
  w_.counts.ensure(m);
  w_.offsets.ensure(m);
  w_.partials.ensure(scan_partial_entries(m));
  w_.sort_scratch.ensure_zeroed(onesweep_status_words((uint32_t)m, 32));
}

void Renderer::ensure_target() {
  const size_t px = (size_t)width_ * height_;
  w_.img32.ensure(px);
  w_.img8.ensure(px);
  const uint32_t tiles = ((width_ + kTile - 1) / kTile) * ((height_ + kTile - 1) / kTile);
  w_.ranges.ensure(tiles);
}

void Renderer::ensure_pairs(uint32_t capacity, uint32_t tile_bits) {
  w_.pair_keys.ensure(capacity);
  w_.pair_vals.ensure(capacity);
  w_.pair_keys_alt.ensure(capacity);
  w_.pair_vals_alt.ensure(capacity);
  w_.sort_scratch.ensure_zeroed(onesweep_status_words(capacity, tile_bits));
}

void Renderer::enqueue_frame(const GsplatCamera &cam) {
  const uint32_t n = scene_->count;
  ensure_frame_buffers(n);
  ensure_target();
  const FrameParams fp = make_frame_params(cam.position, cam.rotation_xyzw, cam.vertical_fov_radians,
                                           cam.near_plane, cam.far_plane, width_, height_);
  const uint32_t tiles = fp.tiles_x * fp.tiles_y;
  uint32_t tile_bits = 1;
  while ((1u << tile_bits) < tiles) ++tile_bits;
  if (w_.pair_cap == 0) w_.pair_cap = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(8ull * n, 1u << 20), 1u << 30);
  ensure_pairs(w_.pair_cap, tile_bits);
  hipStream_t st = stream_;
  ScanScratch sc{w_.partials.get(), w_.total.get(), w_.partials.capacity()};

  GS_HIP(hipEventRecord(ev_[0], st));
  launch_fill_u32(w_.visible.get(), 0u, 1, st);
  launch_project(scene_->sh_degree, scene_->soa.get(), n, fp, w_.keys.get(), w_.vals.get(), w_.rec.get(),
                 w_.rect.get(), w_.visible.get(), nullptr, st);
  GS_HIP(hipEventRecord(ev_[1], st));

  uint32_t *k = w_.keys.get(), *v = w_.vals.get(), *ka = w_.keys_alt.get(), *va = w_.vals_alt.get();
  onesweep_sort(&k, &v, &ka, &va, n, nullptr, 32, /*descending=*/true, w_.sort_scratch.get(),
                w_.sort_scratch.capacity(), st);
  w_.sorted_vals = v;
  GS_HIP(hipEventRecord(ev_[2], st));

  const uint32_t n1 = std::max<uint32_t>(n, 1);
  launch_tile_counts(v, w_.visible.get(), w_.rect.get(), n1, w_.counts.get(), st);
  exclusive_scan(w_.counts.get(), w_.offsets.get(), n1, sc, st);
  launch_pairs_finalize(w_.visible.get(), w_.total.get(), w_.pair_cap, w_.m_dev.get(), w_.info.get(), st);
  launch_tile_emit(v, w_.visible.get(), w_.rect.get(), w_.offsets.get(), n, fp.tiles_x, w_.pair_cap,
                   w_.pair_keys.get(), w_.pair_vals.get(), st);
  uint32_t *pk = w_.pair_keys.get(), *pv = w_.pair_vals.get(), *pka = w_.pair_keys_alt.get(),
           *pva = w_.pair_vals_alt.get();
  onesweep_sort(&pk, &pv, &pka, &pva, w_.pair_cap, w_.m_dev.get(), tile_bits, /*descending=*/false,
                w_.sort_scratch.get(), w_.sort_scratch.capacity(), st);
  w_.tile_list = pv;
  launch_tile_ranges(pk, w_.m_dev.get(), w_.pair_cap, tiles, w_.ranges.get(), st);
  GS_HIP(hipEventRecord(ev_[3], st));

  launch_raster(raster_variant_, w_.rec.get(), w_.rect.get(), w_.ranges.get(), pv, width_, height_, fp.tiles_x,
                fp.tiles_y, w_.img32.get(), w_.img8.get(), st);
  GS_HIP(hipEventRecord(ev_[4], st));
  GS_HIP(hipMemcpyAsync(pinned_, w_.info.get(), 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  GS_HIP(hipEventRecord(ev_[5], st));

  frame_w_ = width_;
  frame_h_ = height_;
  last_fp_ = fp;
  last_cam_ = cam;
  frame_valid_ = true;
  frame_pending_ = true;
}

bool Renderer::absorb_frame_info() {
  visible_ = scene_ && scene_->count ? pinned_[0] : 0;
  pairs_ = scene_ && scene_->count ? ((uint64_t)pinned_[2] << 32) | pinned_[1] : 0;
  const bool overflow = pinned_[3] != 0 && scene_ && scene_->count;
  if (overflow) {
    const uint64_t want = pairs_ + pairs_ / 4 + 4096;
    if (want > (1ull << 30)) throw Error(Code::Unsupported, "tile pair count exceeds 2^30");
    w_.pair_cap = (uint32_t)want;
  }
  return overflow;
}

void Renderer::render(const GsplatCamera &cam) {
  const auto t0 = std::chrono::steady_clock::now();
  if (!scene_) throw Error(Code::SceneNotLoaded, "scene not loaded");
  if (!camera_valid(cam)) throw Error(Code::InvalidArgument, "invalid camera");
  // Grow the pair buffers early when the previous frame is known to have
  // overflowed them (that frame is re-rendered by wait()).
  if (frame_pending_ && hipEventQuery(ev_[5]) == hipSuccess) {
    if (absorb_frame_info()) {
      ++overflow_rerenders_;
      enqueue_frame(last_cam_);
      wait();
    }
  }
  enqueue_frame(cam);
  host_ms_ = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void Renderer::wait() {
  if (!stream_) return;
  for (int attempt = 0; frame_pending_; ++attempt) {
    GS_HIP(hipEventSynchronize(ev_[5]));
    if (absorb_frame_info()) {
      if (attempt > 4) throw Error(Code::Internal, "tile pair buffers keep overflowing");
      ++overflow_rerenders_;
      enqueue_frame(last_cam_);
      continue;
    }
    float ms[4] = {0, 0, 0, 0}, total = 0;
    for (int s = 0; s < 4; ++s) GS_HIP(hipEventElapsedTime(&ms[s], ev_[s], ev_[s + 1]));
    GS_HIP(hipEventElapsedTime(&total, ev_[0], ev_[4]));
    // FrameStats (core/src/lib.rs:172-180): visible = drawn = V on this path
    // (render-wgpu/src/lib.rs:565-583, 736).
    last_stats_.frame_ms = host_ms_;
    last_stats_.preprocess_ms = ms[0];
    last_stats_.sort_ms = ms[1];
    last_stats_.raster_ms = ms[2] + ms[3];
    last_stats_.visible_count = visible_;
    last_stats_.drawn_count = visible_;
    last_times_.project_ms = ms[0];
    last_times_.sort_ms = ms[1];
    last_times_.bin_ms = ms[2];
    last_times_.raster_ms = ms[3];
    last_times_.total_ms = total;
    last_times_.splat_count = scene_ ? scene_->count : 0;
    last_times_.visible_count = visible_;
    last_times_.tiles_x = (frame_w_ + kTile - 1) / kTile;
    last_times_.tiles_y = (frame_h_ + kTile - 1) / kTile;
    last_times_.tile_pairs = pairs_;
    frame_pending_ = false;
  }
  GS_HIP(hipStreamSynchronize(stream_));
}

GsplatStats Renderer::stats() {
  wait();
  return last_stats_;
}

GsplatStageTimes Renderer::stage_times() {
  wait();
  return last_times_;
}

void Renderer::readback_rgba8(uint8_t *out, uint64_t len) {
  if (!frame_valid_) throw Error(Code::SceneNotLoaded, "no rendered frame to read back");
  wait();
  const uint64_t need = (uint64_t)frame_w_ * frame_h_ * 4;
  if (len < need) throw Error(Code::InvalidArgument, "output buffer smaller than width*height*4 bytes");
  GS_HIP(hipMemcpyAsync(out, w_.img8.get(), need, hipMemcpyDeviceToHost, stream_));
  GS_HIP(hipStreamSynchronize(stream_));
}

void Renderer::readback_rgba32f(float *out, uint64_t len) {
  if (!frame_valid_) throw Error(Code::SceneNotLoaded, "no rendered frame to read back");
  wait();
  const uint64_t need = (uint64_t)frame_w_ * frame_h_ * 4;
  if (len < need) throw Error(Code::InvalidArgument, "output buffer smaller than width*height*4 floats");
  GS_HIP(hipMemcpyAsync(out, w_.img32.get(), need * sizeof(float), hipMemcpyDeviceToHost, stream_));
  GS_HIP(hipStreamSynchronize(stream_));
}

void Renderer::read_depth_keys(uint32_t *out, uint64_t cap) {
  if (!frame_valid_) throw Error(Code::SceneNotLoaded, "no rendered frame");
  const uint32_t n = scene_->count;
  if (cap < n) throw Error(Code::InvalidArgument, "output buffer smaller than the splat count");
  if (n == 0) return;
  // The depth sort permutes the key buffers in place, so the keys in splat
  // order are regenerated by the same projection kernel for the last camera.
  DevBuf<uint32_t> keys;
  DevBuf<float> tmp;
  keys.ensure(n);
  tmp.ensure(12ull * n);
  launch_project(scene_->sh_degree, scene_->soa.get(), n, last_fp_, keys.get(), nullptr, nullptr, nullptr, nullptr,
                 tmp.get(), stream_);
  GS_HIP(hipMemcpyAsync(out, keys.get(), (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
  GS_HIP(hipStreamSynchronize(stream_));
}

void Renderer::read_sorted(uint32_t *out, uint64_t cap, uint32_t *count) {
  if (!frame_valid_) throw Error(Code::SceneNotLoaded, "no rendered frame");
  wait();
  if (cap < visible_) throw Error(Code::InvalidArgument, "output buffer smaller than the visible count");
  if (visible_)
    GS_HIP(hipMemcpyAsync(out, w_.sorted_vals, (size_t)visible_ * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
  GS_HIP(hipStreamSynchronize(stream_));
  *count = visible_;
}

void Renderer::read_projection(const GsplatCamera &cam, float *out, uint64_t cap) {
  if (!scene_) throw Error(Code::SceneNotLoaded, "scene not loaded");
  if (!camera_valid(cam)) throw Error(Code::InvalidArgument, "invalid camera");
  const uint32_t n = scene_->count;
  if (cap < 12ull * n) throw Error(Code::InvalidArgument, "output buffer smaller than 12 floats per splat");
  if (n == 0) return;
  DevBuf<float> tmp;
  tmp.ensure(12ull * n);
  const FrameParams fp = make_frame_params(cam.position, cam.rotation_xyzw, cam.vertical_fov_radians,
                                           cam.near_plane, cam.far_plane, width_, height_);
  launch_project(scene_->sh_degree, scene_->soa.get(), n, fp, nullptr, nullptr, nullptr, nullptr, nullptr,
                 tmp.get(), stream_);
  GS_HIP(hipMemcpyAsync(out, tmp.get(), 12ull * n * sizeof(float), hipMemcpyDeviceToHost, stream_));
  GS_HIP(hipStreamSynchronize(stream_));
}

void Renderer::read_tile_counts(uint32_t *out, uint64_t cap, uint32_t *tx, uint32_t *ty) {
  if (!frame_valid_) throw Error(Code::SceneNotLoaded, "no rendered frame");
  wait();
  const uint32_t txs = (frame_w_ + kTile - 1) / kTile, tys = (frame_h_ + kTile - 1) / kTile;
  const size_t tiles = (size_t)txs * tys;
  if (cap < tiles) throw Error(Code::InvalidArgument, "output buffer smaller than the tile count");
  std::vector<uint2> r(tiles);
  GS_HIP(hipMemcpyAsync(r.data(), w_.ranges.get(), tiles * sizeof(uint2), hipMemcpyDeviceToHost, stream_));
  GS_HIP(hipStreamSynchronize(stream_));
  for (size_t t = 0; t < tiles; ++t) out[t] = r[t].y - r[t].x;
  *tx = txs;
  *ty = tys;
}

}  // namespace gs
