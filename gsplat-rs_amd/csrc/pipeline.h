// pipeline.h -- device buffers, resident scenes and the batched frame
// pipeline of the SortedAlpha path on one GPU.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "scratch_state.h"

namespace gs {

#define GS_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      (void)hipGetLastError(); /* a later launch check must not see this error */     \
      throw Error(Code::Internal, std::string(#call) + ": " + hipGetErrorString(e_)); \
    }                                                                                  \
  } while (0)

// Grow-only device allocation; ensure() returns true when it (re)allocated.
// Without a stream, releasing waits for the device, since work queued on any
// stream may still use the old allocation. A buffer grown on a stream (the
// only stream whose work uses it) is freed and made again in that stream's
// order (hipFreeAsync / hipMallocAsync): work queued before still sees the
// old block, and no other stream waits, so one server pipeline growing its
// buffers does not stall the batches of the others. Such a buffer must be
// released (release()) while its stream exists.
template <typename T>
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  // At least `count` elements; a stream-ordered growth takes `headroom`
  // times the count (>= 1), so a slowly growing need reallocates rarely.
  bool ensure(size_t count, hipStream_t stream = nullptr, double headroom = 1.0) {
    if (count <= cap_) return false;
    release();
    size_t want = std::max<size_t>(count, 1);
    if (stream && headroom > 1.0) want = std::max(want, (size_t)((double)want * headroom));
    const hipError_t e = stream ? hipMallocAsync(reinterpret_cast<void **>(&ptr_), want * sizeof(T), stream)
                                : hipMalloc(reinterpret_cast<void **>(&ptr_), want * sizeof(T));
    if (e != hipSuccess) {
      ptr_ = nullptr;
      (void)hipGetLastError();
      throw Error(Code::Internal, "device allocation of " + std::to_string(want * sizeof(T)) +
                                      " bytes failed: " + hipGetErrorString(e));
    }
    cap_ = want;
    owner_ = stream;
    return true;
  }
  // As ensure(), clearing the allocation to zero whenever it is (re)made;
  // true when it was. The clear is queued on `stream`, the stream whose work
  // reads the buffer next: a null-stream hipMemset is not ordered before work
  // on a non-blocking stream. Never called under graph capture.
  bool ensure_zeroed(size_t count, hipStream_t stream, double headroom = 1.0) {
    if (!ensure(count, stream, headroom)) return false;
    GS_HIP(hipMemsetAsync(ptr_, 0, cap_ * sizeof(T), stream));
    return true;
  }
  void release() {
    if (ptr_) {
      if (owner_) {
        (void)hipFreeAsync(ptr_, owner_);
      } else {
        (void)hipDeviceSynchronize();
        (void)hipFree(ptr_);
      }
    }
    ptr_ = nullptr;
    cap_ = 0;
    owner_ = nullptr;
  }
  T *get() const { return ptr_; }
  size_t capacity() const { return cap_; }

 private:
  T *ptr_ = nullptr;
  size_t cap_ = 0;
  hipStream_t owner_ = nullptr;  // the stream a stream-ordered allocation belongs to
};

bool device_present();

// Scene resident in HBM in the coefficient-major SoA layout of
// prepare_scene(); shared by contexts, batches and the server.
struct DeviceScene {
  uint32_t count = 0;
  uint32_t sh_degree = 0;
  // Geometry path (gsplat.h GsplatGeometryPath): 0 the fp32 SoA of
  // prepare_scene, 1 the packed records of pack_scene followed by their
  // decode constants.
  uint32_t geometry_path = 0;
  DevBuf<float> soa;
  // Host statistics for launch choices: centroid of the positions, RMS
  // distance from it, and the median over splats of the largest axis scale
  // exp(max log-scale).
  float center[3] = {0.0f, 0.0f, 0.0f};
  float radius = 0.0f;
  float median_scale = 0.0f;
  bool packed() const { return geometry_path == 1; }
  uint32_t format() const { return scene_format(sh_degree, packed()); }
  uint64_t bytes() const {
    return packed() ? ((uint64_t)packed_words(sh_degree) * count + kPackedConsts) * sizeof(uint32_t)
                    : (uint64_t)soa_fields(sh_degree) * count * sizeof(float);
  }
};
// Uploads `s` in geometry path `geometry_path` (0 direct, 1 packed; packed
// refuses SH degree 4 with Unsupported).
std::shared_ptr<DeviceScene> upload_scene(const HostScene &s, hipStream_t stream, uint32_t geometry_path = 0);

// Device bytes a batch frame needs beyond its scenes: per-slot buffers,
// targets, tile ranges and the initial tile-pair buffers.
uint64_t frame_working_bytes(uint64_t slots, uint64_t pixels, uint64_t tiles, uint64_t pair_cap);

// One view of a batch: a resident scene rendered with the constants `fp`
// (camera and target size).
struct ViewSpec {
  std::shared_ptr<DeviceScene> scene;
  FrameParams fp;
};

// Stage timings of the last frame (device events), in milliseconds.
struct FrameTimes {
  float project_ms = 0, sort_ms = 0, bin_ms = 0, raster_ms = 0, total_ms = 0;
  uint64_t blend_evals = 0;  // counting mode only: evaluations passing coverage
  uint64_t blend_tests = 0;  // counting mode only: coverage tests made
};

// Launch configuration of a pipeline. Every setting gives bit-identical
// images; they trade work between binning, sorting and blending.
struct Tuning {
  uint32_t bin_shift = 2;                                // binning tiles of 16 << bin_shift pixels (the
                                                         // largest, with bin_auto)
  uint32_t bin_auto = 1;                                 // per view: smaller binning tiles for splats
                                                         // that are small on screen
  float bin_auto_k = 3.0f;                               // binning tile edge in splat diameters (bin_auto)
  uint32_t bin_max_ctiles = 0;                           // bin_auto, batches: at most this many binning
                                                         // tiles per view where a shift up to 3 allows
                                                         // (0: no bound; the server uses 512)
  RasterVariant raster = RasterVariant::Unrolled;        // blend loop
  uint32_t rows2_min_tiles = 24576;                      // Unrolled frames of >= this many 16x16 tiles, all
                                                         // views at one bin shift >= 1, at most one splat
                                                         // slot per two pixels, blend two rows per thread
                                                         // (Rows2; 0: never): fewer LDS reads and shared
                                                         // column terms pay once the chip is full
  uint32_t rts_min_tiles = 256;                          // radix sorts of >= this many tiles: reduce-then-scan
  uint32_t blend_min = 64;                               // staged splats that start a blend (1..256)
};

// One candidate measured by Pipeline::autotune.
struct TuneSample {
  Tuning tuning;
  float frame_ms;  // median device time of the frame
};

// Renders batches of views on its own stream. Per frame (Renderer::
// render_frame, render-wgpu/src/lib.rs:723-742, for every view at once):
//   project   depth key + cull + projection + SH colour, one thread per slot
//   sort      one stable radix sort of (depth key, view<<24|index) over all
//             slots, descending depth, then stably by view: every view's
//             segment ends up back to front, ties by splat index
//   bin       per-slot tile counts, scan, (global tile, slot) pairs front to
//             back, stable radix sort by global tile id, tile ranges
//   raster    one workgroup per 16x16 tile of every view
// No host synchronisation inside a frame: the visible counts and the pair
// count are produced and consumed on the device. The pair buffers have a
// host-chosen capacity; a frame whose pair count exceeds it drops the excess,
// flags itself, and is rendered again by wait() after the buffers grow.
class Pipeline {
 public:
  Pipeline();
  ~Pipeline();
  Pipeline(const Pipeline &) = delete;
  Pipeline &operator=(const Pipeline &) = delete;

  // Enqueues one frame of `views` (1..kMaxViews) and returns without waiting
  // for the device. The scenes stay referenced until the next enqueue. When
  // `host_rgba8` is given, entry v (if not null) receives view v's RGBA8
  // image: the raster writes it there directly when the memory is pinned and
  // device-visible (hipHostMalloc), else a copy queued behind the raster
  // does. It must stay valid until wait() returns. Such a view keeps no
  // device image (readback_* of it reads stale data).
  // With `reuse_order` (one view only, has_order() true) the frame skips the
  // depth sort and bins the view in the last kept order (see set_keep_order).
  void enqueue(std::vector<ViewSpec> views, std::vector<uint8_t *> host_rgba8 = {}, bool reuse_order = false);
  // Keep each one-view frame's depth order (a device copy of the sorted keys
  // and values) for later frames of the same scene that reuse it.
  void set_keep_order(bool on) { keep_order_ = on; }
  // A kept order of `scene` exists.
  bool has_order(const DeviceScene *scene) const;
  // Waits for the last frame, re-rendering it after an overflow.
  void wait();
  bool has_frame() const { return !views_.empty(); }

  uint32_t view_count() const { return (uint32_t)views_.size(); }
  const ViewDev &view(uint32_t v) const;
  uint32_t visible(uint32_t v);  // waits
  uint64_t pairs();              // (16x16 tile, splat) pairs; waits
  uint64_t bin_pairs();          // (binning tile, splat) pairs emitted and sorted; waits
  uint64_t total_splats() const { return total_splats_; }
  uint32_t total_tiles() const { return total_tiles_; }
  FrameTimes times();  // waits
  uint64_t overflow_rerenders() const { return overflow_rerenders_.load(std::memory_order_relaxed); }
  // Frames that had to grow (reallocate) any of the frame buffers.
  uint64_t buffer_grows() const { return buffer_grows_.load(std::memory_order_relaxed); }
  // Frames whose bucket depth sort overflowed (rendered again with the radix
  // passes; the next 64 frames use them too).
  uint64_t sort_overflows() const { return sort_overflows_; }
  // Depth sorts of one-view frames of at most kBucketMaxKeys slots in
  // look-back mode as a bucket sort (default on; also GSPLAT_SORT_BUCKET=0/1).
  // The order is the same either way.
  void set_bucket_sort(bool on) { bucket_sort_ = on; }
  // How the last frame's depth sort ran: 0 look-back radix passes, 1
  // reduce-then-scan radix passes, 2 bucket sort, 3 none (order reused).
  uint32_t depth_sort_mode() const { return depth_sort_mode_; }
  uint32_t pair_capacity() const { return pair_cap_; }

  void readback_rgba8(uint32_t v, uint8_t *out, uint64_t len);
  void readback_rgba32f(uint32_t v, float *out, uint64_t len);
  // Back-to-front splat indices of view v (its visible set).
  void read_sorted(uint32_t v, uint32_t *out, uint64_t cap, uint32_t *count);
  void read_tile_counts(uint32_t v, uint32_t *out, uint64_t cap, uint32_t *tx, uint32_t *ty);
  // Depth keys (keys_out, n entries) and/or the 12-float projection export
  // (proj_out, 12n entries) of `scene` under `fp`, into host memory.
  void export_projection(const DeviceScene &scene, const FrameParams &fp, uint32_t *keys_out, float *proj_out);

  void set_raster_variant(RasterVariant v) { tuning_.raster = v; }
  // Graph mode (also GSPLAT_GRAPH=1): each frame is replayed from a hipGraph
  // captured for the batch's configuration (views, sizes, scenes, buffers,
  // tuning); only the view table (cameras) is uploaded per frame. Frames with
  // host RGBA8 targets always launch directly.
  void set_use_graph(bool on) { use_graph_ = on; }
  bool use_graph() const { return use_graph_; }
  // Counting mode (also GSPLAT_COUNT_EVALS=1): the raster counts the
  // (pixel, splat) evaluations that pass coverage before each pixel's early
  // stop; the count is reported with the frame times.
  void set_count_evals(bool on) { count_evals_ = on; }
  // Per-stage device times for the frames that follow (events at each stage
  // boundary, which cost the device a drain each); otherwise only the total.
  void set_stage_timing(bool on) { stage_timing_ = on; }
  bool stage_timing() const { return stage_timing_; }
  // Binning tiles of (16 << shift) pixels, shift <= 3 (also GSPLAT_BIN_SHIFT);
  // images and tile counts do not depend on it.
  void set_bin_shift(uint32_t shift);
  uint32_t bin_shift() const { return tuning_.bin_shift; }
  void set_tuning(const Tuning &t);
  const Tuning &tuning() const { return tuning_; }
  // Renders `views` `reps` times under each candidate tuning (bin shift 0..3
  // x sort mode threshold {128, 256, 1024}), keeps the fastest by median
  // device frame time and returns every measurement. The last frame rendered
  // is under the kept tuning.
  std::vector<TuneSample> autotune(const std::vector<ViewSpec> &views, uint32_t reps);
  hipStream_t stream() const { return stream_; }

 private:
  struct FramePlan {
    int slot = 0;  // pinned view-table staging slot
    uint32_t nviews = 0, slots = 0, tiles = 0, ctiles = 0, tile_bits = 0, count_blocks = 0, hint = 0;
    bool uniform_shift = false;  // every view at one bin shift: heaviest-first raster order
    int value_byte = -1;
    bool reuse_order = false, keep_order = false;
    bool depth_lookback = false, pair_lookback = false;
    bool depth_bucket = false;  // the depth sort is a bucket sort (bucket_sort, sort.h)
    size_t depth_dirty = 0, pair_dirty = 0;                // look-back words the frame clears after use
    size_t depth_rts_words = 0, pair_rts_words = 0;        // words a reduce-then-scan sort may write
    ScratchState::Plan depth_scratch, pair_scratch;
  };
  struct GraphSlot {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> key;
    uint32_t *sorted_keys = nullptr, *sorted_vals = nullptr;
    ScratchState depth_state, pair_state;  // the scratch state a replay leaves
  };
  void launch_frame();
  // Every stream operation of a frame; `captured` leaves out the event
  // records (an event recorded under capture cannot be timed or waited on).
  void record_frame(const FramePlan &plan, bool captured);
  void drop_graph(GraphSlot &g);
  bool absorb_summary();  // true when the frame's pairs overflowed
  void check_view(uint32_t v) const;

  // Declared first, so destroyed last: the buffers below that grew in the
  // stream's order free themselves on it when they are destroyed.
  struct StreamOwner {
    hipStream_t s = nullptr;
    ~StreamOwner() {
      if (s) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
      }
    }
  } stream_owner_;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev_[6] = {};        // stage boundaries 0..4, 5 = summary copied
  hipEvent_t table_ev_[2] = {};  // view-table uploads out of each pinned slot
  uint8_t *table_pinned_[2] = {};
  uint32_t *summary_pinned_ = nullptr;  // [-, M lo, M hi, overflow, visible[kMaxViews]]
  uint64_t frame_counter_ = 0;

  DevBuf<uint32_t> order_;   // heaviest-first binning tile order (kOrderClasses x binning tiles)
  DevBuf<uint8_t> table_;    // seg_start | tile_start | ViewDev | ProjGroup, packed per frame
  DevBuf<uint32_t> summary_;  // device side of summary_pinned_
  DevBuf<uint32_t> keys_, vals_, keys_alt_, vals_alt_;
  DevBuf<float4> rec_;
  DevBuf<uint2> rect_;
  DevBuf<uint32_t> counts_, offsets_;
  DevBuf<uint32_t> pair_keys_, pair_vals_, pair_keys_alt_, pair_vals_alt_;
  DevBuf<uint2> ranges_;
  DevBuf<float4> img32_;
  DevBuf<uchar4> img8_;
  DevBuf<uint32_t> depth_scratch_, pair_scratch_, m_dev_, block_sums_, block_pairs_;
  DevBuf<ViewDev> export_view_;
  DevBuf<uint32_t> order_keys_, order_vals_;  // kept depth order (set_keep_order)
  DevBuf<uint2> key_range_;                   // per projection workgroup key ranges (bucket sort)
  const DeviceScene *order_scene_ = nullptr;  // the scene whose order is kept
  uint32_t order_count_ = 0;
  bool keep_order_ = false, reuse_order_ = false;

  std::vector<ViewSpec> views_;
  std::vector<uint8_t *> host_rgba8_;
  std::vector<ViewDev> host_views_;
  std::vector<ProjGroup> host_groups_;                 // batches only (empty for one view)
  std::pair<uint32_t, uint32_t> group_chunks_[kFormats] = {};  // per scene format: first chunk, chunk count
  uint32_t total_slots_ = 0, total_tiles_ = 0, total_ctiles_ = 0, total_rblocks_ = 0;
  uint64_t total_pixels_ = 0, total_splats_ = 0;
  uint32_t pair_cap_ = 0;
  uint32_t *sorted_keys_ = nullptr;  // depth-sorted keys of the last frame
  uint32_t *sorted_vals_ = nullptr;  // depth-sorted values of the last frame
  bool pending_ = false;
  bool summary_valid_ = false;
  std::vector<uint32_t> visible_;
  uint64_t pairs_ = 0;         // binning pairs of the last frame
  uint64_t raster_pairs_ = 0;  // (16x16 tile, splat) pairs of the last frame
  std::atomic<uint64_t> overflow_rerenders_{0};  // read by other threads (server statistics)
  std::atomic<uint64_t> buffer_grows_{0};        // likewise
  bool bucket_sort_ = true;
  uint32_t bucket_off_ = 0;      // frames left that sort by radix passes after a bucket overflow
  uint64_t sort_overflows_ = 0;
  uint32_t depth_sort_mode_ = 0;
  FrameTimes times_{};
  Tuning tuning_{};
  bool count_evals_ = false;
  ScratchState depth_state_, pair_state_;  // what the depth and pair sorts' scratch holds
  bool use_graph_ = false;
  GraphSlot graphs_[2];
  bool staged_times_ = false;  // the last frame recorded its stage events
  bool stage_timing_ = false;  // record stage events (set_stage_timing, GSPLAT_STAGE_TIMES)

};

}  // namespace gs
