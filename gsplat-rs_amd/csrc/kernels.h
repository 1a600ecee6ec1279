// kernels.h -- host launchers of the gfx950 kernels of the SortedAlpha path.
//
// Every per-frame kernel works on a batch of views described by a device
// view table. A view renders one resident scene from one camera into its own
// target. A single context frame is a batch of one view.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gs {

// Per-view constants of project_splat (splat_surface_direct.wgsl:131-217),
// computed on the host from the camera and target size
// (make_surface_render_params, render-wgpu/src/lib.rs:889-920).
struct FrameParams {
  float cam[3];
  float rows[9];  // quat_to_mat3(quat_inverse(q)), row-major
  float near_plane, far_plane;
  float f, aspect, fx, fy;
  float lim_x, lim_y;
  float blur_x, blur_y;
  uint32_t width, height;
  uint32_t tiles_x, tiles_y;    // 16x16 raster tiles
  uint32_t ctiles_x, ctiles_y;  // binning tiles of (16 << bin shift) pixels
};

// ctiles_x/ctiles_y are left for the caller to set for its bin shift
// (set_bin_tiles).
FrameParams make_frame_params(const float position[3], const float rotation_xyzw[4], float fov,
                              float near_plane, float far_plane, uint32_t width, uint32_t height);

constexpr uint32_t kTile = 16;            // raster tile edge in pixels
constexpr uint32_t kMaxBinShift = 3;      // binning tiles of 16, 32, 64 or 128 pixels
constexpr uint32_t kThreads = 256;        // threads per workgroup (4 wave64)
constexpr uint32_t kCountBlock = kThreads * 4;  // slots per tile-count workgroup
constexpr uint32_t kMaxGrid = 4096;       // bound of the grids that loop over device-sized work
constexpr uint32_t kSegmentAlign = 4096;  // view slot segments align to the sort tile
constexpr uint32_t kLocalBits = 24;       // splat index bits in a sort value
constexpr uint32_t kLocalMask = (1u << kLocalBits) - 1;
constexpr uint32_t kMaxViews = 256;       // view id fits the value's top byte
constexpr uint32_t kXcds = 8;             // MI355X: workgroup b is dispatched to XCD b % 8
constexpr uint32_t kPairViewShift = 24;   // tile-pair key = view << 24 | tile of the view
constexpr uint32_t kPairTileMask = (1u << kPairViewShift) - 1;

// Scene format of a view: an fp32 SoA scene of SH degree d is format d
// (0..4); a packed scene (common.h pack_scene) of SH degree d is
// kPackedFormat + d (d = 0..3). Its device image at `soa` is then the word
// array followed by the kPackedConsts decode constants.
constexpr uint32_t kPackedFormat = 5;
constexpr uint32_t kFormats = kPackedFormat + 4;
inline uint32_t scene_format(uint32_t degree, bool packed) { return packed ? kPackedFormat + degree : degree; }

// One view of a batch. Its splats own the slot segment [seg_off,
// seg_off + seg_len) of every per-splat array (seg_len = n rounded up to
// kSegmentAlign), its tiles own the global tile ids [tile_off, tile_off +
// tiles), and its target starts at pixel img_off.
struct ViewDev {
  FrameParams fp;
  const float *soa;  // scene image (coefficient-major, see common.h)
  uint32_t n, format;
  uint32_t seg_off, seg_len;
  uint32_t tile_off, tiles;    // global 16x16 raster tiles
  uint32_t ctile_off, ctiles;  // global binning tiles
  uint32_t bin_shift;          // binning tiles of (16 << bin_shift) pixels
  uint32_t rblock_off;         // first raster workgroup (see launch_raster)
  uint64_t img_off;
  // Caller's pinned host target (device-visible): the raster writes the
  // view's RGBA8 pixels there and keeps no device image of it. Null: the
  // device images at img_off.
  uchar4 *out8;
};

// Binning tile grid of a view for bin shift `shift`.
void set_bin_tiles(FrameParams &fp, uint32_t shift);

// Raster workgroups of a view (binning tiles in whole runs of kXcds, 4^shift
// raster tiles each).
inline uint32_t raster_blocks(uint32_t ctiles, uint32_t shift) {
  return (ctiles + kXcds - 1) / kXcds * kXcds << (2 * shift);
}

// Projected splat, 3 x float4 per slot (48 B):
//   [cx, cy, i00, i01] [i10, i11, cut, alpha] [r*a, g*a, b*a, 0]
// where (cx, cy) is the NDC centre, i = [axis_u axis_v]^-1, cut the coverage
// bound on r^2 and (r,g,b)*a the premultiplied colour.
constexpr uint32_t kRecordFloat4 = 3;

// Up to three device word ranges a kernel zeroes with its whole grid before
// its own work (ranges the frame needs zero later on, folded into an earlier
// launch).
struct Clears {
  static constexpr int kRanges = 3;
  uint32_t *ptr[kRanges] = {nullptr, nullptr, nullptr};
  size_t words[kRanges] = {0, 0, 0};
  void add(uint32_t *p, size_t n) {
    for (int r = 0; r < kRanges; ++r)
      if (words[r] == 0) {
        ptr[r] = p;
        words[r] = n;
        return;
      }
  }
};

// Depth key, cull, projection, SH colour and pixel bounds for every slot of
// every view whose scene has scene format `format` (other views are skipped):
// keys[g] = f32 bits of depth when near <= depth <= far else 0;
// vals[g] = view << 24 | splat index (index 0xFFFFFF for padding slots);
// rect[g] = pixel bounds x0|y0<<16, x1|y1<<16, or x = ~0u when the splat
// covers no pixel.
// The grid first zeroes `clears`. With `key_range` (total_slots / 256
// entries), workgroup b (slots [256 b, 256 b + 256)) also stores its smallest
// non-zero key (~0u when none) and its largest key at key_range[b], for a
// bucket sort of the keys (sort.h).
// With `table` (one view), the view table arrives by value with the launch
// instead of through a device copy: the projection works from it and its
// first workgroup stores it at table->dst for the kernels after it.
constexpr uint32_t kInlineTableWords = 64;
constexpr uint32_t kInlineViewWord = 8;  // word offset of the view's ViewDev in a one-view table
struct InlineTable {
  alignas(16) uint32_t words[kInlineTableWords];
  uint32_t count = 0;      // words of the table
  uint32_t view_word = 0;  // word offset of the view's ViewDev (kInlineViewWord)
  uint32_t *dst = nullptr;
};
void launch_project(uint32_t format, const ViewDev *views, const uint32_t *seg_start, uint32_t nviews,
                    uint32_t total_slots, uint32_t *keys, uint32_t *vals, float4 *rec, uint2 *rect,
                    const Clears &clears, hipStream_t stream, uint2 *key_range = nullptr,
                    const InlineTable *table = nullptr);

// Views of one scene projected together (views of a batch that share a
// scene, at most kGroupViews per group). Group chunks are runs of 256
// splats: [chunk_start, chunk_start + seg_len / 256), covering the padding
// slots of the views' segments too; groups are ordered by chunk_start.
constexpr uint32_t kGroupViews = 8;
struct ProjGroup {
  const float *soa;  // the scene (all views of the group)
  uint32_t n, format;
  uint32_t chunk_start, nv;
  uint32_t views[kGroupViews];
};

// As launch_project for the groups whose chunks are [chunk_lo, chunk_lo +
// chunks) (all of scene format `format`): each splat's scene fields are read
// (and a packed record decoded) once for all views of its group.
void launch_project_groups(uint32_t format, const ViewDev *views, const ProjGroup *groups, uint32_t ngroups,
                           uint32_t chunk_lo, uint32_t chunks, uint32_t *keys, uint32_t *vals, float4 *rec,
                           uint2 *rect, const Clears &clears, hipStream_t stream);

// The same projection of one view's n splats exporting, per splat, the depth
// key (keys_out, optional) and the reference quantities of project_splat +
// eval_color (12 floats: drawn, centre, axis_u, axis_v, premultiplied rgba).
void launch_project_export(uint32_t format, const ViewDev *views, uint32_t view, uint32_t n, uint32_t *keys_out,
                           float *export12, hipStream_t stream);

// Scratch of the two-level exclusive scan.
// Two-level binning. Splats are binned into binning tiles of (16 << shift)
// pixels; each 16x16 raster tile then keeps, in order, the splats of its
// binning tile whose pixel bounds overlap it. shift 0 bins straight into the
// raster tiles. Either way a raster tile blends exactly the splats whose
// bounds overlap it, in depth order.

// After the depth sort (each view's segment back to front, culled last),
// over total_slots / kCountBlock blocks of reversed slot positions
// (total_slots a multiple of kCountBlock):
//   counts_rev[total-1-g] = number of binning tiles of the splat in sorted
//   slot g (0 for culled and padding slots), offsets_rev its exclusive scan
//   within the block, block_sums[b] the block's total and block_pairs[b] its
//   16x16-tile pair count;
//   visible[v] = V of view v (written by the last visible slot; zero first);
//   rect_lo[g], rect_hi[g] = the pixel bounds of sorted slot g's splat (only
//   meaningful where counts_rev is non-zero), for the emission.
// The grid first zeroes `clears`.
void launch_tile_counts(const ViewDev *views, const uint32_t *seg_start, uint32_t nviews, uint32_t total_slots,
                        const uint32_t *skeys, const uint32_t *svals, const uint2 *rect,
                        uint32_t *counts_rev, uint32_t *offsets_rev, uint32_t *visible, uint32_t *block_sums,
                        uint32_t *block_pairs, uint32_t *rect_lo, uint32_t *rect_hi, const Clears &clears,
                        hipStream_t stream);

// One workgroup: block_sums becomes its exclusive scan (each block's first
// pair index, so a slot's first pair is block_sums[r / kCountBlock] +
// offsets_rev[r]); info[1..2] = the 64-bit binning pair total M, info[3] = 1
// when M exceeds `capacity`, *m_out = min(M, capacity); *raster_pairs = the
// sum of block_pairs.
void launch_block_scan(uint32_t *block_sums, uint32_t blocks, uint32_t capacity, uint32_t *m_out, uint32_t *info,
                       const uint32_t *block_pairs, unsigned long long *raster_pairs, hipStream_t stream);

// Digit histograms of a look-back pair sort of at most kPairHistPasses
// passes, accumulated by the pair emission instead of a histogram launch
// (sort.h SortPrep::hist_ready): hist[(stripe * kSortMaxPasses + p) *
// 2^digit_bits + digit] with digit (key >> shift[p]) & (2^digit_bits - 1)
// and stripe = emission workgroup mod kSortStripes; hist zero beforehand.
constexpr uint32_t kPairHistPasses = 2;
struct PairHist {
  uint32_t *hist = nullptr;  // null: no histogram
  uint32_t passes = 0, digit_bits = 8;
  uint32_t shift[kPairHistPasses] = {0, 0};
};

// Writes (view << 24 | binning tile of the view, record slot) pairs at the
// scanned offsets: each view's pairs form one contiguous run, front to back.
// Pairs at or beyond `capacity` are dropped (the frame is then flagged and
// rendered again). With hist.hist, also counts the written pairs' digits.
// With totals.blocks (the count blocks, at most kEmitScanMaxBlocks),
// block_sums holds the blocks' raw totals (launch_block_scan has not run):
// each workgroup sums the totals before its own block itself, and
// workgroup 0 writes what the block scan would have (info[1..3], *m_out,
// *raster_pairs), one launch fewer for frames of few slots.
constexpr uint32_t kEmitScanMaxBlocks = 1024;
struct EmitTotals {
  uint32_t blocks = 0;
  const uint32_t *block_pairs = nullptr;
  uint32_t *m_out = nullptr;
  uint32_t *info = nullptr;
  unsigned long long *raster_pairs = nullptr;
};
void launch_tile_emit(const ViewDev *views, const uint32_t *seg_start, uint32_t nviews, uint32_t total_slots,
                      const uint32_t *svals, const uint32_t *rect_lo, const uint32_t *rect_hi,
                      const uint32_t *counts_rev,
                      const uint32_t *offsets_rev, const uint32_t *block_sums, uint32_t capacity,
                      uint32_t *pair_keys, uint32_t *pair_vals, hipStream_t stream, const PairHist &hist = PairHist(),
                      const EmitTotals &totals = EmitTotals());

// Words copied by one workgroup of a kernel, dst[i] = src[i] for i < words
// (dst may be pinned host memory: the frame's summary, read by the host
// after the frame's last event, without a copy launch of its own).
struct WordCopy {
  const uint32_t *src = nullptr;
  uint32_t *dst = nullptr;
  uint32_t words = 0;
};

// After a stable sort of the pairs on the tile bits of their keys:
// ranges[t] = [first, last+1) of global binning tile t (= ctile_off of the
// key's view + the key's tile; the ranges must be zero beforehand); the pair
// count is read from *m_dev (grid sized for `capacity`). The grid first
// zeroes `clears`; workgroup 0 also performs `copy`.
void launch_tile_ranges(const ViewDev *views, const uint32_t *pair_keys, const uint32_t *m_dev, uint32_t capacity,
                        uint2 *ranges, const Clears &clears, const WordCopy &copy, hipStream_t stream);

// 16x16-tile counts of view v (row-major, tiles_x * tiles_y, zeroed here):
// the number of visible splats whose pixel bounds overlap each raster tile.
void launch_raster_tile_counts(const ViewDev *views, uint32_t v, const ViewDev &host_view, const uint32_t *skeys,
                               const uint32_t *svals, const uint2 *rect, uint32_t *counts, hipStream_t stream);

// Heaviest-first order of the batch's binning tiles for the raster: eight
// classes by list length (class 0 longest), order[class * total_ctiles +
// place] = global binning tile, counts[class] = class size (zero on entry).
constexpr uint32_t kOrderClasses = 8;
// Class of a binning tile with `len` list entries: clamp(12 - floor(log2
// len), 0, kOrderClasses - 1) (class 0: 4096 or more; empty lists last).
__host__ __device__ inline uint32_t raster_order_class(uint32_t len) {
  const int lg = len ? 31 - __builtin_clz(len) : 0;
  return (uint32_t)(lg >= 12 ? 0 : (12 - lg < (int)kOrderClasses ? 12 - lg : (int)kOrderClasses - 1));
}
void launch_raster_order(const uint2 *ranges, uint32_t total_ctiles, uint32_t *counts, uint32_t *order,
                         hipStream_t stream);

// Raster workgroup layout: order null for per-view ranges (rblock_start);
// else the heaviest-first order, every view at bin shift `shift`.
struct RasterOrder {
  const uint32_t *order = nullptr;
  const uint32_t *counts = nullptr;
  uint32_t total_ctiles = 0;
  uint32_t shift = 0;
};

// Front-to-back blend of every view's 16x16 tiles: one workgroup per tile,
// the raster tiles of one binning tile consecutive on one XCD. View v owns
// the workgroups [rblock_start[v], rblock_start[v] + raster_blocks(view v))
// (raster_blocks: the view's binning tiles rounded up to whole kXcds runs,
// times 4^bin_shift); total_rblocks is their sum over the batch.
// The workgroup walks its binning tile's list 256 entries at a time, keeps
// (in order) the splats whose bounds overlap its tile, stages their records
// in LDS and blends them once 256 are staged or the list ends.
//   Linear    per-lane exit at the early stop, divergent blend branch
//   Unrolled  wave-uniform splat loop, four splats per wave liveness check
//             (default)
//   Rows2     as Unrolled with two pixels per thread (two tile rows per
//             workgroup, one LDS read of a splat for both pixels); every
//             view's bin shift must be >= 1 (both rows in one binning tile)
//   Queue     as Unrolled in a persistent grid (about the workgroups
//             resident at once) that takes the (view, tile) work items from
//             queues in the same order and XCD placement; `queue`
//             (kRasterQueueWords words) must be zero at launch
//   Blocks8   8x8 raster tiles: each wave of a 16x16 workgroup blends one
//             8x8 block over the splats whose pixel bounds meet it (the
//             records staged once per workgroup, a position list per block)
// All give bit-identical images. With `evals`, the counting variant also
// adds to evals[0] the number of (pixel, splat) evaluations that pass
// coverage before each pixel's early stop (the blend's algorithmic work) and
// to evals[1] the coverage tests made. `clears` and `copy` are work the
// launch does for the kernels before it (see raster_prologue): words to zero
// and the frame summary to hand to pinned memory; neither may overlap what
// the raster reads.
enum class RasterVariant : uint32_t { Linear = 0, Unrolled = 1, Rows2 = 2, Queue = 3, Blocks8 = 4 };
constexpr uint32_t kQueueLanes = 16;   // persistent raster: queues per XCD
constexpr uint32_t kQueueStride = 16;  // words between two queue counters (their own cache lines)
constexpr uint32_t kRasterQueueWords = kXcds * kQueueLanes * kQueueStride;
void launch_raster(RasterVariant variant, const ViewDev *views, const uint32_t *rblock_start, uint32_t nviews,
                   uint32_t total_rblocks, const RasterOrder &order, uint32_t blend_min, const float4 *rec, const uint2 *rect, const uint2 *ranges,
                   const uint32_t *list, float4 *img32, uchar4 *img8, hipStream_t stream,
                   unsigned long long *evals = nullptr, uint32_t *queue = nullptr, const Clears &clears = Clears(),
                   const WordCopy &copy = WordCopy());

// Device fill of n u32 words (graph-safe replacement of hipMemsetAsync).
void launch_fill_u32(uint32_t *p, uint32_t value, size_t n, hipStream_t stream);

}  // namespace gs
