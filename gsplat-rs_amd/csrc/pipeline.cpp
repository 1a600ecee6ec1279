// pipeline.cpp -- batched frame pipeline of the SortedAlpha path.

#include "pipeline.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <algorithm>

#include "sort.h"

namespace gs {

namespace {

// Device view table of a frame, packed in 16-byte aligned sections:
// seg_start[nviews] | rblock_start[nviews] | ViewDev[nviews] | ProjGroup[ngroups]
// (first slot and first raster workgroup of each view).
struct TableLayout {
  size_t rblock_start, views, groups, bytes;
};
constexpr size_t align16(size_t x) { return (x + 15) & ~size_t(15); }
constexpr TableLayout table_layout(uint32_t nviews, uint32_t ngroups) {
  const size_t rblock_start = align16((size_t)nviews * sizeof(uint32_t));
  const size_t views = align16(rblock_start + (size_t)nviews * sizeof(uint32_t));
  const size_t groups = align16(views + (size_t)nviews * sizeof(ViewDev));
  return TableLayout{rblock_start, views, groups, groups + (size_t)ngroups * sizeof(ProjGroup)};
}
constexpr size_t kTableBytes = table_layout(kMaxViews, kMaxViews).bytes;
// Frame summary words (device and pinned host copies):
//   [0] bucket depth sort overflow, [1..2] binning pair count M (u64),
//   [3] overflow flag, [4..5] (16x16 tile, splat) pairs (u64),
//   [6 .. 6+kMaxViews) visible count per view,
//   kEvalsWord: u64 evaluations passing, u64 coverage tests (counting mode),
//   kOrderWord: class sizes of the raster's heaviest-first order,
//   kQueueWord: the persistent raster's queue counters.
constexpr uint32_t kInlineTableMaxSlots = 256 * kThreads;  // projection workgroups that take the inline table
constexpr size_t kSortOverflowWord = 0;
constexpr size_t kRasterPairsWord = 4;
constexpr size_t kVisibleWord = 6;
constexpr size_t kEvalsWord = kVisibleWord + kMaxViews;
constexpr size_t kOrderWord = kEvalsWord + 4;  // raster order class sizes (kOrderClasses words)
constexpr size_t kQueueWord = kOrderWord + kOrderClasses;  // persistent raster queues (kRasterQueueWords)
constexpr size_t kSummaryWords = kQueueWord + kRasterQueueWords;

static_assert(kSegmentAlign % kSortTile == 0, "view segments must be whole sort tiles");

uint32_t bits_for(uint64_t count) {
  uint32_t b = 1;
  while (b < 32 && (1ull << b) < count) ++b;
  return b;
}

uint64_t div_up_u64(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

uint32_t initial_pair_capacity(uint64_t splats) {
  return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(8 * splats, 1u << 20), 1u << 30);
}

// Binning tile of a view: about `bin_auto_k` splat diameters, no larger
// than the tuning's. A splat's on-screen radius is estimated as 3 sigma of
// the scene's median largest scale at the depth of the scene's centroid
// (at least half the scene's RMS radius away). Splats that are small next to
// the binning tile make each 16x16 tile scan a list mostly of splats outside
// it; large ones multiply the pairs of small binning tiles. In a batch, with
// bin_max_ctiles set, a view is then binned coarser (up to 128 px) until it
// has at most that many binning tiles: with 512 or fewer the pair sort is one
// pass (9-bit digits). That pays for batches of many small views (served
// requests) and costs batches of large views with large splats (their raster
// lists grow more than one pass saves), so it is off by default.
uint32_t view_bin_shift(const DeviceScene &sc, const FrameParams &fp, const Tuning &t, bool batch) {
  const uint32_t lo = t.raster == RasterVariant::Rows2 ? 1u : 0u;  // two tile rows share a list
  if (!t.bin_auto || !(sc.median_scale > 0.0f)) return std::max(t.bin_shift, lo);
  const float rel[3] = {sc.center[0] - fp.cam[0], sc.center[1] - fp.cam[1], sc.center[2] - fp.cam[2]};
  const float z = std::max(fp.rows[6] * rel[0] + fp.rows[7] * rel[1] + fp.rows[8] * rel[2],
                           std::max(0.5f * sc.radius, 1e-3f));
  const float r_px = 3.0f * sc.median_scale * fp.f * 0.5f * (float)fp.height / z;
  const float tile = t.bin_auto_k * 2.0f * r_px;
  uint32_t s = 0;
  while (s < t.bin_shift && (float)(kTile << (s + 1)) <= tile) ++s;
  s = std::max(s, lo);
  if (batch && t.bin_max_ctiles) {  // coarser until the view has at most bin_max_ctiles binning tiles
    auto ctiles = [&](uint32_t sh) {
      const uint32_t e = kTile << sh;
      return ((fp.width + e - 1) / e) * ((fp.height + e - 1) / e);
    };
    while (s < kMaxBinShift && ctiles(s) > t.bin_max_ctiles) ++s;
  }
  return s;
}

}  // namespace

bool device_present() {
  int n = 0;
  return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

std::shared_ptr<DeviceScene> upload_scene(const HostScene &s, hipStream_t stream, uint32_t geometry_path) {
  if (geometry_path > 1) throw Error(Code::Unsupported, "geometry path " + std::to_string(geometry_path) + " is not built");
  std::vector<float> image;
  auto d = std::make_shared<DeviceScene>();
  d->geometry_path = geometry_path;
  if (geometry_path == 1) {
    PackedScene p = pack_scene(s);
    d->count = p.count;
    d->sh_degree = p.sh_degree;
    image.resize(p.words.size() + p.consts.size());
    std::memcpy(image.data(), p.words.data(), p.words.size() * sizeof(uint32_t));
    std::memcpy(image.data() + p.words.size(), p.consts.data(), p.consts.size() * sizeof(float));
  } else {
    PreparedScene p = prepare_scene(s);
    d->count = p.count;
    d->sh_degree = p.sh_degree;
    image = std::move(p.soa);
  }
  if (s.count > 0 && s.positions.size() >= 3 * (size_t)s.count && s.scale_xyz.size() >= 3 * (size_t)s.count) {
    double c[3] = {0.0, 0.0, 0.0};
    for (uint32_t i = 0; i < s.count; ++i)
      for (int a = 0; a < 3; ++a) c[a] += s.positions[3 * (size_t)i + a];
    for (int a = 0; a < 3; ++a) d->center[a] = (float)(c[a] / s.count);
    double r2 = 0.0;
    std::vector<float> ls(s.count);
    for (uint32_t i = 0; i < s.count; ++i) {
      for (int a = 0; a < 3; ++a) {
        const double q = s.positions[3 * (size_t)i + a] - d->center[a];
        r2 += q * q;
      }
      ls[i] = std::max(s.scale_xyz[3 * (size_t)i], std::max(s.scale_xyz[3 * (size_t)i + 1], s.scale_xyz[3 * (size_t)i + 2]));
    }
    d->radius = (float)std::sqrt(r2 / s.count);
    std::nth_element(ls.begin(), ls.begin() + ls.size() / 2, ls.end());
    d->median_scale = std::exp(ls[ls.size() / 2]);
  }
  d->soa.ensure(image.size());
  if (d->count > 0) {
    GS_HIP(hipMemcpyAsync(d->soa.get(), image.data(), image.size() * sizeof(float), hipMemcpyHostToDevice, stream));
    GS_HIP(hipStreamSynchronize(stream));
  }
  return d;
}

uint64_t frame_working_bytes(uint64_t slots, uint64_t pixels, uint64_t tiles, uint64_t pair_cap) {
  // keys, vals and their alternates, record (48 B), rect, counts, offsets.
  const uint64_t per_slot = 4 * 4 + 16 * kRecordFloat4 + 8 + 4 + 4;
  // pair keys/vals and alternates, plus the sort status words of the larger
  // of the two sorts.
  const uint64_t sort_words = onesweep_status_words((uint32_t)std::max<uint64_t>(slots, pair_cap), 32);
  // targets, tile ranges and the raster order (bounded by the 16x16 tiles).
  return per_slot * slots + 20 * pixels + 8 * tiles + 4 * kOrderClasses * tiles + 16 * pair_cap + 4 * sort_words;
}

Pipeline::Pipeline() {
  if (!device_present()) throw Error(Code::Unsupported, "gpu rasterizer unavailable");
  if (const char *e = std::getenv("GSPLAT_COUNT_EVALS")) count_evals_ = std::atoi(e) != 0;
  if (const char *e = std::getenv("GSPLAT_GRAPH")) use_graph_ = std::atoi(e) != 0;
  if (const char *e = std::getenv("GSPLAT_STAGE_TIMES")) stage_timing_ = std::atoi(e) != 0;
  if (const char *e = std::getenv("GSPLAT_BIN_SHIFT")) set_bin_shift((uint32_t)std::atoi(e));
  if (const char *e = std::getenv("GSPLAT_BIN_AUTO")) tuning_.bin_auto = std::atoi(e) != 0;
  if (const char *e = std::getenv("GSPLAT_BIN_MAX_CTILES")) tuning_.bin_max_ctiles = (uint32_t)std::strtoul(e, nullptr, 10);
  if (const char *e = std::getenv("GSPLAT_BIN_AUTO_K")) {
    const float k = std::strtof(e, nullptr);
    if (k > 0.0f) tuning_.bin_auto_k = k;
  }
  if (const char *e = std::getenv("GSPLAT_BLEND_MIN"))
    tuning_.blend_min = std::min<uint32_t>(std::max<uint32_t>((uint32_t)std::strtoul(e, nullptr, 10), 1u), 256u);
  if (const char *e = std::getenv("GSPLAT_RASTER_ROWS2_MIN_TILES"))
    tuning_.rows2_min_tiles = (uint32_t)std::strtoul(e, nullptr, 10);
  if (const char *e = std::getenv("GSPLAT_SORT_BUCKET")) bucket_sort_ = std::atoi(e) != 0;
  if (const char *e = std::getenv("GSPLAT_SORT_RTS_MIN_TILES"))
    tuning_.rts_min_tiles = std::max<uint32_t>((uint32_t)std::strtoul(e, nullptr, 10), 1u);
  GS_HIP(hipSetDevice(0));
  GS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  stream_owner_.s = stream_;
  // The frame-start and stage-boundary events only time the device: no
  // system-scope fence (a cache writeback and invalidation per event). The
  // frame's last event keeps it: the host then reads what the frame wrote.
  for (int e = 0; e < 6; ++e)
    GS_HIP(hipEventCreateWithFlags(&ev_[e], e < 5 ? hipEventDisableSystemFence : hipEventDefault));
  for (int r = 0; r < 2; ++r) {
    GS_HIP(hipEventCreateWithFlags(&table_ev_[r], hipEventDisableTiming));
    GS_HIP(hipEventRecord(table_ev_[r], stream_));
    GS_HIP(hipHostMalloc(reinterpret_cast<void **>(&table_pinned_[r]), kTableBytes, hipHostMallocDefault));
  }
  GS_HIP(hipHostMalloc(reinterpret_cast<void **>(&summary_pinned_), kSummaryWords * sizeof(uint32_t),
                       hipHostMallocDefault));
  std::memset(summary_pinned_, 0, kSummaryWords * sizeof(uint32_t));
  table_.ensure(kTableBytes);
  summary_.ensure(kSummaryWords);
  m_dev_.ensure(1);
  export_view_.ensure(1);
}

Pipeline::~Pipeline() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (GraphSlot &g : graphs_) drop_graph(g);
  for (auto &e : ev_)
    if (e) (void)hipEventDestroy(e);
  for (int r = 0; r < 2; ++r) {
    if (table_ev_[r]) (void)hipEventDestroy(table_ev_[r]);
    if (table_pinned_[r]) (void)hipHostFree(table_pinned_[r]);
  }
  if (summary_pinned_) (void)hipHostFree(summary_pinned_);
  // the buffers free on stream_, which stream_owner_ destroys after them
}

bool Pipeline::has_order(const DeviceScene *scene) const {
  return scene && order_scene_ == scene && order_count_ == scene->count && order_keys_.get();
}

void Pipeline::enqueue(std::vector<ViewSpec> views, std::vector<uint8_t *> host_rgba8, bool reuse_order) {
  if (views.empty()) throw Error(Code::InvalidArgument, "batch has no views");
  if (reuse_order && (views.size() != 1 || !has_order(views[0].scene.get())))
    throw Error(Code::Internal, "no kept depth order for this view");
  if (!host_rgba8.empty() && host_rgba8.size() != views.size())
    throw Error(Code::InvalidArgument, "one host target per view expected");
  if (views.size() > kMaxViews)
    throw Error(Code::InvalidArgument, "batch of " + std::to_string(views.size()) + " views exceeds " +
                                           std::to_string(kMaxViews));
  // A finished earlier frame that overflowed only grows the pair buffers:
  // its results are replaced by this frame.
  if (pending_) {
    const hipError_t q = hipEventQuery(ev_[5]);
    if (q == hipSuccess) {
      absorb_summary();
      pending_ = false;
    } else {
      (void)hipGetLastError();  // "not ready" is no error of the next launch
      if (q != hipErrorNotReady) GS_HIP(q);
    }
  }
  const bool multi = views.size() > 1;
  std::vector<ViewDev> table(views.size());
  uint64_t slots = 0, tiles = 0, ctiles = 0, rblocks = 0, pixels = 0, splats = 0;
  for (size_t v = 0; v < views.size(); ++v) {
    const ViewSpec &s = views[v];
    if (!s.scene) throw Error(Code::SceneNotLoaded, "scene not loaded");
    const uint32_t n = s.scene->count;
    if (multi && n > kLocalMask)
      throw Error(Code::Unsupported, "a scene in a batch of several views holds at most " +
                                         std::to_string(kLocalMask) + " splats");
    ViewDev &d = table[v];
    d.fp = s.fp;
    d.bin_shift = view_bin_shift(*s.scene, s.fp, tuning_, multi);
    set_bin_tiles(d.fp, d.bin_shift);
    d.soa = s.scene->soa.get();
    d.n = n;
    d.format = s.scene->format();
    d.seg_off = (uint32_t)slots;
    d.seg_len = (uint32_t)(((uint64_t)n + kSegmentAlign - 1) / kSegmentAlign * kSegmentAlign);
    d.tile_off = (uint32_t)tiles;
    d.tiles = d.fp.tiles_x * d.fp.tiles_y;
    d.ctile_off = (uint32_t)ctiles;
    d.ctiles = d.fp.ctiles_x * d.fp.ctiles_y;
    d.rblock_off = (uint32_t)rblocks;
    rblocks += raster_blocks(d.ctiles, d.bin_shift);
    d.img_off = pixels;
    d.out8 = nullptr;
    if (!host_rgba8.empty() && host_rgba8[v]) {
      // A pinned host target the raster can write directly; otherwise the
      // image is copied there after the frame.
      void *dev = nullptr;
      if (hipHostGetDevicePointer(&dev, host_rgba8[v], 0) == hipSuccess && dev)
        d.out8 = reinterpret_cast<uchar4 *>(dev);
      else
        (void)hipGetLastError();
    }
    slots += d.seg_len;
    tiles += d.tiles;
    ctiles += d.ctiles;
    pixels += (uint64_t)s.fp.width * s.fp.height;
    splats += n;
    if (slots > (1ull << 30)) throw Error(Code::Unsupported, "batch exceeds 2^30 splat slots");
  }
  if (tiles >= (1ull << 30) || rblocks >= (1ull << 31)) throw Error(Code::Unsupported, "batch exceeds 2^30 tiles");

  // Projection groups of a batch: views of one scene (at most kGroupViews
  // each) project together, reading the scene once. Groups are ordered by
  // scene format, so each format's chunks form one range (one launch).
  host_groups_.clear();
  for (auto &r : group_chunks_) r = {0, 0};
  if (multi) {
    uint32_t chunk = 0;
    for (uint32_t fmt = 0; fmt < kFormats; ++fmt) {
      const size_t first = host_groups_.size();
      group_chunks_[fmt].first = chunk;
      for (uint32_t v = 0; v < (uint32_t)table.size(); ++v) {
        const ViewDev &d = table[v];
        if (d.format != fmt) continue;
        ProjGroup *g = nullptr;
        for (size_t q = first; q < host_groups_.size() && !g; ++q)
          if (host_groups_[q].soa == d.soa && host_groups_[q].n == d.n && host_groups_[q].nv < kGroupViews)
            g = &host_groups_[q];
        if (!g) {
          ProjGroup ng{};
          ng.soa = d.soa;
          ng.n = d.n;
          ng.format = fmt;
          host_groups_.push_back(ng);
          g = &host_groups_.back();
        }
        g->views[g->nv++] = v;
      }
      for (size_t q = first; q < host_groups_.size(); ++q) {
        host_groups_[q].chunk_start = chunk;
        chunk += table[host_groups_[q].views[0]].seg_len / kThreads;
      }
      group_chunks_[fmt].second = chunk - group_chunks_[fmt].first;
    }
  }

  reuse_order_ = reuse_order;
  views_ = std::move(views);
  host_rgba8_ = std::move(host_rgba8);
  host_views_ = std::move(table);
  total_slots_ = (uint32_t)slots;
  total_tiles_ = (uint32_t)tiles;
  total_ctiles_ = (uint32_t)ctiles;
  total_rblocks_ = (uint32_t)rblocks;
  total_pixels_ = pixels;
  total_splats_ = splats;
  if (pair_cap_ == 0) pair_cap_ = initial_pair_capacity(splats);
  launch_frame();
}

void Pipeline::launch_frame() {
  const uint32_t nviews = (uint32_t)views_.size();
  const uint32_t slots = total_slots_, tiles = total_tiles_;
  const uint32_t ctiles = total_ctiles_;
  uint32_t view_ctiles = 1;
  for (const ViewDev &d : host_views_) view_ctiles = std::max(view_ctiles, d.ctiles);
  const uint32_t tile_bits = bits_for(view_ctiles);  // the pair sort orders binning tiles within a view
  const int value_byte = nviews > 1 ? 3 : -1;
  const size_t m = std::max<uint32_t>(slots, 1);
  const uint32_t count_blocks = slots / kCountBlock;
  // Buffers grow in the stream's order with 25% headroom (no other stream
  // waits; a batch mix that varies reallocates rarely).
  const hipStream_t st = stream_;
  constexpr double kHead = 1.25;
  bool grew = false;
  grew |= keys_.ensure(m, st, kHead);
  grew |= vals_.ensure(m, st, kHead);
  grew |= keys_alt_.ensure(m, st, kHead);
  grew |= vals_alt_.ensure(m, st, kHead);
  grew |= rec_.ensure(kRecordFloat4 * m, st, kHead);
  grew |= rect_.ensure(m, st, kHead);
  grew |= counts_.ensure(m, st, kHead);
  grew |= offsets_.ensure(m, st, kHead);
  grew |= block_sums_.ensure(count_blocks + 1, st, kHead);
  grew |= block_pairs_.ensure(count_blocks + 1, st, kHead);
  grew |= img32_.ensure(total_pixels_, st, kHead);
  grew |= img8_.ensure(total_pixels_, st, kHead);
  grew |= ranges_.ensure(std::max<uint32_t>(ctiles, 1), st, kHead);
  grew |= order_.ensure((size_t)kOrderClasses * std::max<uint32_t>(ctiles, 1), st, kHead);
  grew |= pair_keys_.ensure(pair_cap_, st);
  grew |= pair_vals_.ensure(pair_cap_, st);
  grew |= pair_keys_alt_.ensure(pair_cap_, st);
  grew |= pair_vals_alt_.ensure(pair_cap_, st);
  // Look-back sorts need zeroed scratch; a later kernel of the frame clears
  // what each sort dirtied, so the next frame finds it clean.
  if (depth_scratch_.ensure_zeroed(std::max(onesweep_status_words((uint32_t)m, 32), bucket_sort_words((uint32_t)m)),
                                   st, kHead)) {
    depth_state_.reset();
    grew = true;
  }
  if (pair_scratch_.ensure_zeroed(onesweep_status_words(pair_cap_, tile_bits), st)) {
    pair_state_.reset();
    grew = true;
  }
  if (grew) buffer_grows_.fetch_add(1, std::memory_order_relaxed);
  const bool depth_lookback = radix_uses_lookback((uint32_t)m, 0, tuning_.rts_min_tiles);
  // One short view sorts its depth keys by bucket (four launches instead of
  // five, one pass over the keys instead of four); after a bucket overflow
  // the next 64 frames use the radix passes.
  const bool depth_bucket =
      depth_lookback && nviews == 1 && !reuse_order_ && bucket_sort_ && bucket_off_ == 0 && m <= kBucketMaxKeys;
  if (bucket_off_) --bucket_off_;
  if (depth_bucket && key_range_.ensure(m / kThreads + 1, st, kHead)) buffer_grows_.fetch_add(1, std::memory_order_relaxed);
  // The last frame's pair count (+25%) sizes the pair sort's grids and mode.
  uint32_t hint = pairs_ ? (uint32_t)std::min<uint64_t>(pairs_ + pairs_ / 4, pair_cap_) : pair_cap_;
  // A graph replays the grids it was captured with: round the hint up to a
  // power of two sort tiles so that moving cameras re-capture rarely.
  const bool as_graph = use_graph_ && host_rgba8_.empty();
  if (as_graph)
    hint = (uint32_t)std::min<uint64_t>((uint64_t)kSortTile << bits_for(div_up_u64(hint, kSortTile)), pair_cap_);
  const bool pair_lookback = radix_uses_lookback(pair_cap_, hint, tuning_.rts_min_tiles);
  const RadixPasses depth_ps = radix_passes(32, true, value_byte), pair_ps = radix_passes(tile_bits, false, -1);
  // A bucket sort dirties its look-back words and the range words after them.
  const size_t depth_dirty = depth_bucket ? bucket_sort_words((uint32_t)m) : radix_lookback_words((uint32_t)m, depth_ps);
  const size_t pair_dirty = radix_lookback_words(pair_cap_, pair_ps);

  // View table through a pinned staging slot whose previous upload is done.
  const int r = (int)(frame_counter_++ % 2);
  GS_HIP(hipEventSynchronize(table_ev_[r]));
  uint8_t *host = table_pinned_[r];
  const TableLayout lay = table_layout(nviews, (uint32_t)host_groups_.size());
  uint32_t *seg_start = reinterpret_cast<uint32_t *>(host);
  uint32_t *rblock_start = reinterpret_cast<uint32_t *>(host + lay.rblock_start);
  for (uint32_t v = 0; v < nviews; ++v) {
    seg_start[v] = host_views_[v].seg_off;
    rblock_start[v] = host_views_[v].rblock_off;
  }
  std::memcpy(host + lay.views, host_views_.data(), nviews * sizeof(ViewDev));
  if (!host_groups_.empty())
    std::memcpy(host + lay.groups, host_groups_.data(), host_groups_.size() * sizeof(ProjGroup));

  FramePlan plan;
  plan.slot = r;
  plan.nviews = nviews;
  plan.slots = slots;
  plan.tiles = tiles;
  plan.ctiles = ctiles;
  plan.tile_bits = tile_bits;
  plan.value_byte = value_byte;
  plan.count_blocks = count_blocks;
  plan.hint = hint;
  plan.depth_lookback = depth_lookback;
  plan.depth_bucket = depth_bucket;
  depth_sort_mode_ = reuse_order_ ? 3u : depth_bucket ? 2u : depth_lookback ? 0u : 1u;
  plan.pair_lookback = pair_lookback;
  plan.depth_dirty = depth_dirty;
  plan.pair_dirty = pair_dirty;
  plan.depth_scratch = depth_state_.before(depth_lookback, depth_dirty);
  plan.pair_scratch = pair_state_.before(pair_lookback, pair_dirty);
  plan.depth_rts_words = radix_rts_words((uint32_t)m, depth_ps);
  plan.pair_rts_words = radix_rts_words(pair_cap_, pair_ps);
  plan.reuse_order = reuse_order_;
  plan.keep_order = keep_order_ && nviews == 1 && !reuse_order_;
  if (plan.keep_order) {
    order_keys_.ensure(m, st, kHead);
    order_vals_.ensure(m, st, kHead);
    order_scene_ = views_[0].scene.get();
    order_count_ = views_[0].scene->count;
  }
  plan.uniform_shift = true;
  for (const ViewDev &d : host_views_) plan.uniform_shift = plan.uniform_shift && d.bin_shift == host_views_[0].bin_shift;
  if (as_graph) {
    // The frame as one graph per staging slot, captured again whenever
    // anything but the view table's contents changes.
    std::vector<uint64_t> key = {(uint64_t)r, nviews, slots, tiles, ctiles, tile_bits, (uint64_t)(value_byte + 1),
                                 pair_cap_, hint, tuning_.bin_shift,
                                 (uint64_t)tuning_.raster, tuning_.rows2_min_tiles, tuning_.rts_min_tiles, tuning_.blend_min,
                                 tuning_.bin_auto, count_evals_, depth_state_.rts_dirty(), pair_state_.rts_dirty(),
                                 (uint64_t)reuse_order_, (uint64_t)keep_order_, (uint64_t)depth_bucket};
    for (const ViewDev &d : host_views_) {
      key.push_back(reinterpret_cast<uintptr_t>(d.soa));
      key.push_back(((uint64_t)d.n << 8) | d.format);
      key.push_back(((uint64_t)d.fp.width << 32) | d.fp.height);
      key.push_back(d.bin_shift);
    }
    for (const void *ptr : {(const void *)keys_.get(), (const void *)vals_.get(), (const void *)keys_alt_.get(),
                            (const void *)vals_alt_.get(), (const void *)rec_.get(), (const void *)rect_.get(),
                            (const void *)counts_.get(), (const void *)offsets_.get(),
                            (const void *)block_sums_.get(), (const void *)block_pairs_.get(),
                            (const void *)img32_.get(), (const void *)img8_.get(), (const void *)ranges_.get(),
                            (const void *)pair_keys_.get(), (const void *)pair_vals_.get(),
                            (const void *)pair_keys_alt_.get(), (const void *)pair_vals_alt_.get(),
                            (const void *)depth_scratch_.get(), (const void *)pair_scratch_.get(),
                            (const void *)order_.get()})
      key.push_back(reinterpret_cast<uintptr_t>(ptr));
    GraphSlot &g = graphs_[r];
    if (!g.exec || g.key != key) {
      drop_graph(g);
      GS_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
      try {
        record_frame(plan, true);
      } catch (...) {
        hipGraph_t partial = nullptr;
        (void)hipStreamEndCapture(stream_, &partial);
        if (partial) (void)hipGraphDestroy(partial);
        throw;
      }
      GS_HIP(hipStreamEndCapture(stream_, &g.graph));
      GS_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
      g.key = std::move(key);
      g.sorted_keys = sorted_keys_;
      g.sorted_vals = sorted_vals_;
      g.depth_state = depth_state_;
      g.pair_state = pair_state_;
    }
    GS_HIP(hipEventRecord(ev_[0], stream_));
    GS_HIP(hipGraphLaunch(g.exec, stream_));
    for (int e = 1; e <= 4; ++e) GS_HIP(hipEventRecord(ev_[e], stream_));
    GS_HIP(hipEventRecord(table_ev_[r], stream_));
    GS_HIP(hipEventRecord(ev_[5], stream_));
    staged_times_ = false;
    sorted_keys_ = g.sorted_keys;
    sorted_vals_ = g.sorted_vals;
    depth_state_ = g.depth_state;
    pair_state_ = g.pair_state;
  } else {
    record_frame(plan, false);
    staged_times_ = stage_timing_;
  }
  pending_ = true;
  summary_valid_ = false;
}

void Pipeline::drop_graph(GraphSlot &g) {
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (g.graph) (void)hipGraphDestroy(g.graph);
  g.exec = nullptr;
  g.graph = nullptr;
  g.key.clear();
}

void Pipeline::record_frame(const FramePlan &plan, bool captured) {
  const uint32_t nviews = plan.nviews, slots = plan.slots, ctiles = plan.ctiles;
  const uint32_t tile_bits = plan.tile_bits, count_blocks = plan.count_blocks, hint = plan.hint;
  const int value_byte = plan.value_byte, r = plan.slot;
  const bool depth_lookback = plan.depth_lookback, pair_lookback = plan.pair_lookback;
  const size_t depth_dirty = plan.depth_dirty, pair_dirty = plan.pair_dirty;
  hipStream_t st = stream_;
  const uint32_t ngroups = (uint32_t)host_groups_.size();
  const TableLayout lay = table_layout(nviews, ngroups);
  // Events break the back-to-back run of kernels (the device drains at each,
  // about 6 us measured): a frame records its start and end, and the stage
  // boundaries only with stage timing on.
  const bool stage_events = !captured && stage_timing_;
  if (!captured) GS_HIP(hipEventRecord(ev_[0], st));
  // A launched (not captured) frame of one short view hands its view table
  // to the projection by value, which stores it for the later kernels: no
  // upload launch (a captured graph must copy it from the staging slot on
  // replay). Measured: 640x480 16k splats 7172 -> 7407 frames/s, but 1080p
  // 300k 4930 -> 4870 (every one of its 1172 projection workgroups then reads
  // the view from the launch arguments), hence the limit of 256 workgroups.
  const bool inline_table = nviews == 1 && ngroups == 0 && !captured && slots > 0 &&
                            slots <= kInlineTableMaxSlots && lay.bytes <= kInlineTableWords * sizeof(uint32_t);
  InlineTable inline_tab;
  if (inline_table) {
    std::memcpy(inline_tab.words, table_pinned_[r], lay.bytes);
    inline_tab.count = (uint32_t)((lay.bytes + 3) / 4);
    inline_tab.view_word = (uint32_t)(lay.views / 4);
    inline_tab.dst = reinterpret_cast<uint32_t *>(table_.get());
  } else {
    GS_HIP(hipMemcpyAsync(table_.get(), table_pinned_[r], lay.bytes, hipMemcpyHostToDevice, st));
  }
  const uint32_t *seg_start_d = reinterpret_cast<const uint32_t *>(table_.get());
  const uint32_t *rblock_start_d = reinterpret_cast<const uint32_t *>(table_.get() + lay.rblock_start);
  const ViewDev *views_d = reinterpret_cast<const ViewDev *>(table_.get() + lay.views);
  const ProjGroup *groups_d = reinterpret_cast<const ProjGroup *>(table_.get() + lay.groups);
  uint32_t *info = summary_.get();
  uint32_t *visible = info + kVisibleWord;
  unsigned long long *raster_pairs = reinterpret_cast<unsigned long long *>(info + kRasterPairsWord);

  // The projection zeroes the summary words and the binning tile ranges of
  // this frame.
  Clears frame_clears;
  frame_clears.add(info, kVisibleWord + nviews);
  frame_clears.add(reinterpret_cast<uint32_t *>(ranges_.get()), 2 * (size_t)ctiles);
  frame_clears.add(info + kOrderWord, kOrderClasses + kRasterQueueWords);  // order class sizes, raster queues
  if (slots == 0) {
    for (int c = 0; c < Clears::kRanges; ++c) launch_fill_u32(frame_clears.ptr[c], 0u, frame_clears.words[c], st);
  }
  if (ngroups == 0) {
    launch_project(host_views_[0].format, views_d, seg_start_d, nviews, slots, keys_.get(), vals_.get(),
                   rec_.get(), rect_.get(), frame_clears, st, plan.depth_bucket ? key_range_.get() : nullptr,
                   inline_table ? &inline_tab : nullptr);
  } else {
    for (uint32_t fmt = 0; fmt < kFormats; ++fmt)
      if (group_chunks_[fmt].second) {
        launch_project_groups(fmt, views_d, groups_d, ngroups, group_chunks_[fmt].first, group_chunks_[fmt].second,
                              keys_.get(), vals_.get(), rec_.get(), rect_.get(), frame_clears, st);
        frame_clears = Clears();  // once
      }
  }
  if (stage_events) GS_HIP(hipEventRecord(ev_[1], st));

  uint32_t *k = keys_.get(), *v = vals_.get(), *ka = keys_alt_.get(), *va = vals_alt_.get();
  const bool sorted = !plan.reuse_order;
  if (plan.reuse_order) {
    // The last refreshed order, presented with this frame's projection: the
    // splats it holds are binned (those now culled have no pixel bounds),
    // and no depth sort runs. keys_alt_ / vals_alt_ stay free for the rects.
    k = order_keys_.get();
    v = order_vals_.get();
  } else {
    if (plan.depth_scratch.fill) launch_fill_u32(depth_scratch_.get(), 0u, plan.depth_scratch.fill, st);
    if (plan.depth_bucket) {
      // an overflow flags summary word 0; wait() then renders the frame again
      bucket_sort(&k, &v, &ka, &va, slots, key_range_.get(), slots / kThreads, depth_scratch_.get(),
                  depth_scratch_.capacity(), plan.depth_scratch.skip_clear, info + kSortOverflowWord, st);
    } else {
      SortPrep depth_prep;
      depth_prep.scratch_clean = plan.depth_scratch.skip_clear;
      depth_prep.rts_min_tiles = tuning_.rts_min_tiles;
      depth_prep.seg_start = seg_start_d;  // view v's slots: segment v, value byte 3 = v
      depth_prep.segs = nviews;
      onesweep_sort(&k, &v, &ka, &va, slots, nullptr, 32, /*descending=*/true, depth_scratch_.get(),
                    depth_scratch_.capacity(), st, value_byte, 0, depth_prep);
    }
    if (plan.keep_order) {  // kept for the frames that reuse it
      GS_HIP(hipMemcpyAsync(order_keys_.get(), k, (size_t)slots * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
      GS_HIP(hipMemcpyAsync(order_vals_.get(), v, (size_t)slots * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
    }
  }
  sorted_keys_ = k;
  sorted_vals_ = v;
  if (stage_events) GS_HIP(hipEventRecord(ev_[2], st));

  Clears scan_clears;  // the depth sort's scratch, for the next frame
  if (sorted && depth_lookback) scan_clears.add(depth_scratch_.get(), depth_dirty);
  if (sorted) depth_state_.after(depth_lookback, plan.depth_rts_words);
  // The sort's other buffers (ka, va) are free until the next frame: they
  // carry the rects in sorted order from the counts to the emission.
  launch_tile_counts(views_d, seg_start_d, nviews, slots, k, v, rect_.get(), counts_.get(),
                     offsets_.get(), visible, block_sums_.get(), block_pairs_.get(), ka, va, scan_clears, st);
  // Few count blocks: the emission sums the block totals it needs itself
  // (one launch fewer); many: the block scan launch.
  EmitTotals emit_totals;
  if (count_blocks >= 1 && count_blocks <= kEmitScanMaxBlocks)
    emit_totals = EmitTotals{count_blocks, block_pairs_.get(), m_dev_.get(), info, raster_pairs};
  else
    launch_block_scan(block_sums_.get(), count_blocks, pair_cap_, m_dev_.get(), info, block_pairs_.get(),
                      raster_pairs, st);
  // A look-back pair sort of at most two passes takes its digit histograms
  // from the emission (its scratch is clean by then: any fill runs first).
  SortPrep pair_prep;
  if (plan.pair_scratch.fill) launch_fill_u32(pair_scratch_.get(), 0u, plan.pair_scratch.fill, st);
  PairHist pair_hist;
  const RadixPasses pair_ps = radix_passes(tile_bits, /*descending=*/false, -1);
  if (pair_lookback && plan.pair_scratch.skip_clear && pair_ps.count <= kPairHistPasses) {
    pair_hist.hist = pair_scratch_.get();
    pair_hist.passes = pair_ps.count;
    pair_hist.digit_bits = pair_ps.digit_bits;
    for (uint32_t p = 0; p < pair_ps.count; ++p) pair_hist.shift[p] = pair_ps.shift[p];
    pair_prep.hist_ready = true;
  }
  launch_tile_emit(views_d, seg_start_d, nviews, slots, v, ka, va, counts_.get(), offsets_.get(),
                   block_sums_.get(), pair_cap_, pair_keys_.get(), pair_vals_.get(), st, pair_hist, emit_totals);
  uint32_t *pk = pair_keys_.get(), *pv = pair_vals_.get(), *pka = pair_keys_alt_.get(), *pva = pair_vals_alt_.get();
  pair_prep.scratch_clean = plan.pair_scratch.skip_clear;
  pair_prep.rts_min_tiles = tuning_.rts_min_tiles;
  // One view whose pair sort is one look-back pass over emission-fed
  // histograms: the pass also writes the binning tile ranges and the
  // heaviest-first order (the histogram is each tile's list length), and the
  // raster takes over the clears and the summary hand-over, so neither the
  // ranges kernel nor the order kernel runs.
  const bool tiles_from_sort = nviews == 1 && pair_prep.hist_ready && pair_ps.count == 1 && !count_evals_ &&
                               plan.uniform_shift && ctiles <= (1u << pair_ps.digit_bits);
  if (tiles_from_sort) pair_prep.tiles = TileRangesOut{ranges_.get(), ctiles, order_.get(), info + kOrderWord};
  onesweep_sort(&pk, &pv, &pka, &pva, pair_cap_, m_dev_.get(), tile_bits, /*descending=*/false,
                pair_scratch_.get(), pair_scratch_.capacity(), st, -1, hint, pair_prep);
  Clears range_clears;  // the pair sort's scratch (and the counting words), for later
  if (pair_lookback) range_clears.add(pair_scratch_.get(), pair_dirty);
  pair_state_.after(pair_lookback, plan.pair_rts_words);
  if (count_evals_) range_clears.add(info + kEvalsWord, 4);
  // Every summary word but the evaluation counts is final here: the ranges
  // kernel hands them to the pinned summary itself.
  WordCopy summary_copy;
  if (!count_evals_) summary_copy = WordCopy{info, summary_pinned_, (uint32_t)(kVisibleWord + nviews)};
  if (!tiles_from_sort)
    launch_tile_ranges(views_d, pk, m_dev_.get(), pair_cap_, ranges_.get(), range_clears, summary_copy, st);
  // Heaviest binning tiles first when every view bins alike (one layout).
  RasterOrder ro;
  uint32_t rblocks = total_rblocks_;
  if (plan.uniform_shift) {
    if (!tiles_from_sort) launch_raster_order(ranges_.get(), ctiles, info + kOrderWord, order_.get(), st);
    ro.order = order_.get();
    ro.counts = info + kOrderWord;
    ro.total_ctiles = ctiles;
    ro.shift = host_views_[0].bin_shift;
    rblocks = raster_blocks(ctiles, ro.shift);
  }
  if (stage_events) GS_HIP(hipEventRecord(ev_[3], st));

  unsigned long long *evals = count_evals_ ? reinterpret_cast<unsigned long long *>(info + kEvalsWord) : nullptr;
  // Large, sparse frames of one binning layout blend two rows per thread:
  // with the heaviest tiles first the longer workgroups do not stretch the
  // tail. Dense frames (more than one splat slot per two pixels) keep one
  // row: their longest lists set the kernel's length, and two rows per
  // workgroup double it (a batch of 24 served requests: raster 3.7 -> 5.5 ms).
  RasterVariant variant = tuning_.raster;
  if (variant == RasterVariant::Unrolled && plan.uniform_shift && host_views_[0].bin_shift >= 1 &&
      tuning_.rows2_min_tiles && plan.tiles >= tuning_.rows2_min_tiles && 2 * (uint64_t)plan.slots <= total_pixels_)
    variant = RasterVariant::Rows2;
  launch_raster(variant, views_d, rblock_start_d, nviews, rblocks, ro, tuning_.blend_min,
                rec_.get(), rect_.get(),
                ranges_.get(), pv, img32_.get(), img8_.get(), st, evals, info + kQueueWord,
                tiles_from_sort ? range_clears : Clears(), tiles_from_sort ? summary_copy : WordCopy());
  if (stage_events) GS_HIP(hipEventRecord(ev_[4], st));
  for (size_t i = 0; i < host_rgba8_.size(); ++i) {
    if (!host_rgba8_[i] || host_views_[i].out8) continue;  // written by the raster itself
    const ViewDev &d = host_views_[i];
    GS_HIP(hipMemcpyAsync(host_rgba8_[i], img8_.get() + d.img_off, (size_t)d.fp.width * d.fp.height * 4,
                          hipMemcpyDeviceToHost, st));
  }
  if (count_evals_)
    GS_HIP(hipMemcpyAsync(summary_pinned_, info, kSummaryWords * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  if (!captured) {
    GS_HIP(hipEventRecord(ev_[5], st));
    GS_HIP(hipEventRecord(table_ev_[r], st));  // the staging slot may be rewritten
  }
}

bool Pipeline::absorb_summary() {
  const uint32_t nviews = (uint32_t)views_.size();
  visible_.assign(summary_pinned_ + kVisibleWord, summary_pinned_ + kVisibleWord + nviews);
  pairs_ = ((uint64_t)summary_pinned_[2] << 32) | summary_pinned_[1];
  std::memcpy(&raster_pairs_, summary_pinned_ + kRasterPairsWord, sizeof(raster_pairs_));
  const bool overflow = summary_pinned_[3] != 0;
  if (overflow) {
    const uint64_t want = pairs_ + pairs_ / 4 + 4096;
    if (want > (1ull << 30)) throw Error(Code::Unsupported, "tile pair count exceeds 2^30");
    pair_cap_ = (uint32_t)want;
  }
  const bool sort_overflow = summary_pinned_[kSortOverflowWord] != 0;
  if (sort_overflow) {
    bucket_off_ = 64;
    ++sort_overflows_;
  }
  return overflow || sort_overflow;
}

void Pipeline::wait() {
  if (!stream_) return;
  for (int attempt = 0; pending_; ++attempt) {
    GS_HIP(hipEventSynchronize(ev_[5]));
    if (absorb_summary()) {
      if (attempt > 4) throw Error(Code::Internal, "tile pair buffers keep overflowing");
      ++overflow_rerenders_;
      launch_frame();
      continue;
    }
    float ms[4] = {0, 0, 0, 0}, total = 0;
    if (staged_times_) {
      for (int s = 0; s < 4; ++s) GS_HIP(hipEventElapsedTime(&ms[s], ev_[s], ev_[s + 1]));
      GS_HIP(hipEventElapsedTime(&total, ev_[0], ev_[4]));
    } else {
      GS_HIP(hipEventElapsedTime(&total, ev_[0], ev_[5]));
    }
    uint64_t evals[2] = {0, 0};
    if (count_evals_) std::memcpy(evals, summary_pinned_ + kEvalsWord, sizeof(evals));
    times_ = FrameTimes{ms[0], ms[1], ms[2], ms[3], total, evals[0], evals[1]};
    pending_ = false;
    summary_valid_ = true;
  }
  GS_HIP(hipStreamSynchronize(stream_));
}

void Pipeline::check_view(uint32_t v) const {
  if (views_.empty()) throw Error(Code::SceneNotLoaded, "no rendered frame");
  if (v >= views_.size())
    throw Error(Code::InvalidArgument, "view " + std::to_string(v) + " out of range for a batch of " +
                                           std::to_string(views_.size()));
}

const ViewDev &Pipeline::view(uint32_t v) const {
  check_view(v);
  return host_views_[v];
}

uint32_t Pipeline::visible(uint32_t v) {
  check_view(v);
  wait();
  return visible_[v];
}

uint64_t Pipeline::pairs() {
  wait();
  return raster_pairs_;
}

uint64_t Pipeline::bin_pairs() {
  wait();
  return pairs_;
}

void Pipeline::set_bin_shift(uint32_t shift) {
  if (shift > kMaxBinShift) throw Error(Code::InvalidArgument, "bin shift above " + std::to_string(kMaxBinShift));
  tuning_.bin_shift = shift;
}

void Pipeline::set_tuning(const Tuning &t) {
  if (t.bin_shift > kMaxBinShift) throw Error(Code::InvalidArgument, "bin shift above " + std::to_string(kMaxBinShift));
  if (t.raster != RasterVariant::Linear && t.raster != RasterVariant::Unrolled && t.raster != RasterVariant::Rows2 &&
      t.raster != RasterVariant::Queue && t.raster != RasterVariant::Blocks8)
    throw Error(Code::InvalidArgument, "unknown raster variant");
  if (t.rts_min_tiles == 0) throw Error(Code::InvalidArgument, "sort mode threshold must be positive");
  if (t.blend_min == 0 || t.blend_min > 256) throw Error(Code::InvalidArgument, "blend threshold outside 1..256");
  wait();
  tuning_ = t;
}

std::vector<TuneSample> Pipeline::autotune(const std::vector<ViewSpec> &views, uint32_t reps) {
  if (reps == 0) throw Error(Code::InvalidArgument, "autotune needs at least one repetition");
  std::vector<TuneSample> out;
  const Tuning base = tuning_;
  TuneSample best{base, 1e30f};
  auto measure = [&](const Tuning &t) {
    set_tuning(t);
    enqueue(views);  // warm: buffers grow, the pair count settles
    wait();
    std::vector<float> ms;
    for (uint32_t r = 0; r < reps; ++r) {
      enqueue(views);
      wait();
      ms.push_back(times_.total_ms);
    }
    std::sort(ms.begin(), ms.end());
    const TuneSample sample{t, ms[ms.size() / 2]};
    out.push_back(sample);
    if (sample.frame_ms < best.frame_ms) best = sample;
  };
  for (uint32_t shift = 0; shift <= kMaxBinShift; ++shift)
    for (uint32_t rts : {128u, 256u, 1024u})
      for (RasterVariant rv :
           {RasterVariant::Unrolled, RasterVariant::Rows2, RasterVariant::Queue, RasterVariant::Blocks8}) {
      if (rv == RasterVariant::Rows2 && shift == 0) continue;  // runs as Unrolled there
      Tuning t = base;
      t.bin_shift = shift;
      t.rts_min_tiles = rts;
      t.raster = rv;
      t.rows2_min_tiles = 0;  // the variant is chosen here
      measure(t);
    }
  // then the LDS stage depth (blend threshold) around the best of those
  const Tuning grid_best = best.tuning;
  for (uint32_t blend_min : {32u, 128u, 256u}) {
    if (blend_min == grid_best.blend_min) continue;
    Tuning t = grid_best;
    t.blend_min = blend_min;
    measure(t);
  }
  set_tuning(best.tuning);
  enqueue(views);
  wait();
  return out;
}

FrameTimes Pipeline::times() {
  wait();
  return times_;
}

void Pipeline::readback_rgba8(uint32_t v, uint8_t *out, uint64_t len) {
  check_view(v);
  wait();
  const ViewDev &d = host_views_[v];
  const uint64_t need = (uint64_t)d.fp.width * d.fp.height * 4;
  if (len < need) throw Error(Code::InvalidArgument, "output buffer smaller than width*height*4 bytes");
  GS_HIP(hipMemcpyAsync(out, img8_.get() + d.img_off, need, hipMemcpyDeviceToHost, stream_));
  GS_HIP(hipStreamSynchronize(stream_));
}

void Pipeline::readback_rgba32f(uint32_t v, float *out, uint64_t len) {
  check_view(v);
  wait();
  const ViewDev &d = host_views_[v];
  const uint64_t need = (uint64_t)d.fp.width * d.fp.height * 4;
  if (len < need) throw Error(Code::InvalidArgument, "output buffer smaller than width*height*4 floats");
  GS_HIP(hipMemcpyAsync(out, img32_.get() + d.img_off, need * sizeof(float), hipMemcpyDeviceToHost, stream_));
  GS_HIP(hipStreamSynchronize(stream_));
}

void Pipeline::read_sorted(uint32_t v, uint32_t *out, uint64_t cap, uint32_t *count) {
  check_view(v);
  wait();
  const uint32_t vis = visible_[v];
  if (cap < vis) throw Error(Code::InvalidArgument, "output buffer smaller than the visible count");
  const ViewDev &d = host_views_[v];
  if (vis) {
    GS_HIP(hipMemcpyAsync(out, sorted_vals_ + d.seg_off, (size_t)vis * sizeof(uint32_t), hipMemcpyDeviceToHost,
                          stream_));
    GS_HIP(hipStreamSynchronize(stream_));
    const uint32_t base = v << kLocalBits;
    for (uint32_t i = 0; i < vis; ++i) out[i] -= base;
  }
  *count = vis;
}

void Pipeline::read_tile_counts(uint32_t v, uint32_t *out, uint64_t cap, uint32_t *tx, uint32_t *ty) {
  check_view(v);
  wait();
  const ViewDev &d = host_views_[v];
  if (cap < d.tiles) throw Error(Code::InvalidArgument, "output buffer smaller than the tile count");
  const ViewDev *views_d = reinterpret_cast<const ViewDev *>(
      table_.get() + table_layout((uint32_t)host_views_.size(), (uint32_t)host_groups_.size()).views);
  DevBuf<uint32_t> counts;
  counts.ensure(std::max<uint32_t>(d.tiles, 1));
  launch_raster_tile_counts(views_d, v, d, sorted_keys_, sorted_vals_, rect_.get(), counts.get(), stream_);
  if (d.tiles)
    GS_HIP(hipMemcpyAsync(out, counts.get(), d.tiles * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
  GS_HIP(hipStreamSynchronize(stream_));
  *tx = d.fp.tiles_x;
  *ty = d.fp.tiles_y;
}

void Pipeline::export_projection(const DeviceScene &scene, const FrameParams &fp, uint32_t *keys_out,
                                 float *proj_out) {
  const uint32_t n = scene.count;
  if (n == 0) return;
  ViewDev d{};
  d.fp = fp;
  d.soa = scene.soa.get();
  d.n = n;
  d.format = scene.format();
  GS_HIP(hipStreamSynchronize(stream_));
  GS_HIP(hipMemcpy(export_view_.get(), &d, sizeof(d), hipMemcpyHostToDevice));
  DevBuf<uint32_t> keys;
  DevBuf<float> proj;
  keys.ensure(n);
  proj.ensure(12ull * n);
  launch_project_export(scene.format(), export_view_.get(), 0, n, keys.get(), proj.get(), stream_);
  if (keys_out)
    GS_HIP(hipMemcpyAsync(keys_out, keys.get(), (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
  if (proj_out)
    GS_HIP(hipMemcpyAsync(proj_out, proj.get(), 12ull * n * sizeof(float), hipMemcpyDeviceToHost, stream_));
  GS_HIP(hipStreamSynchronize(stream_));
}

}  // namespace gs
