// kernels.h -- host launchers of the gfx950 kernels of the SortedAlpha path.
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
  uint32_t tiles_x, tiles_y;
};

FrameParams make_frame_params(const float position[3], const float rotation_xyzw[4], float fov,
                              float near_plane, float far_plane, uint32_t width, uint32_t height);

constexpr uint32_t kTile = 16;            // raster tile edge in pixels
constexpr uint32_t kThreads = 256;        // threads per workgroup (4 wave64)
constexpr uint32_t kScanTile = kThreads * 16;

// Projected splat, 3 x float4 per splat (48 B):
//   [cx, cy, i00, i01] [i10, i11, cut, alpha] [r*a, g*a, b*a, 0]
// where (cx, cy) is the NDC centre, i = [axis_u axis_v]^-1, cut the coverage
// bound on r^2 and (r,g,b)*a the premultiplied colour.
constexpr uint32_t kRecordFloat4 = 3;

// Projection + depth key + cull + SH colour + tile rectangle for every splat.
// keys[i] = f32 bits of depth if near<=depth<=far else 0; vals[i] = i;
// rect[i] = packed tile rectangle (x0|y0<<16, x1|y1<<16) or x = ~0u when
// the splat covers no pixel; *visible += number of depth-visible splats.
// `export12`, when non-null, receives the reference quantities (12 floats
// per splat) instead of the record/rect outputs.
void launch_project(uint32_t sh_degree, const float *soa, uint32_t n, const FrameParams &fp,
                    uint32_t *keys, uint32_t *vals, float4 *rec, uint2 *rect, uint32_t *visible,
                    float *export12, hipStream_t stream);

// Scratch of the two-level exclusive scan.
struct ScanScratch {
  uint32_t *partials = nullptr;  // first-level partials
  uint64_t *total = nullptr;     // 64-bit grand total (device)
  size_t partial_cap = 0;
};
size_t scan_partial_entries(size_t n);

// Exclusive scan of n u32 values; the 64-bit sum goes to *scratch.total.
void exclusive_scan(const uint32_t *in, uint32_t *out, size_t n, ScanScratch &s, hipStream_t stream);

// counts[j] = number of tiles of the j-th splat front to back, i.e. of
// sorted[V-1-j], for j < V (V read from *visible); 0 for j >= V.
void launch_tile_counts(const uint32_t *sorted, const uint32_t *visible, const uint2 *rect, uint32_t n,
                        uint32_t *counts, hipStream_t stream);

// Frame summary written on the device after the tile-count scan:
// info[0] = V, info[1..2] = 64-bit pair total M, info[3] = 1 when M exceeds
// the pair capacity; *m_out = min(M, capacity) drives the later kernels.
void launch_pairs_finalize(const uint32_t *visible, const uint64_t *total, uint32_t capacity, uint32_t *m_out,
                           uint32_t *info, hipStream_t stream);

// Writes (tile id, splat) pairs front to back at offsets[j]; pairs at or
// beyond `capacity` are dropped (the frame is then flagged and re-rendered).
void launch_tile_emit(const uint32_t *sorted, const uint32_t *visible, const uint2 *rect,
                      const uint32_t *offsets, uint32_t n, uint32_t tiles_x, uint32_t capacity,
                      uint32_t *pair_keys, uint32_t *pair_vals, hipStream_t stream);

// ranges[t] = [first, last+1) of tile t in the tile-sorted pair keys; the
// pair count is read from *m_dev (grid sized for `capacity`).
void launch_tile_ranges(const uint32_t *pair_keys, const uint32_t *m_dev, uint32_t capacity, uint32_t tiles,
                        uint2 *ranges, hipStream_t stream);

// Front-to-back blend of each 16x16 tile's list.
// All variants: one workgroup per tile, batches of 256 records staged in LDS.
//   TileLds      thread = pixel in row-major tile order
//   QuadWave     each wave owns an 8x8 quadrant and skips splats whose pixel
//                bounds miss it
//   QuadWaveXcd  QuadWave with the tiles banded per XCD
enum class RasterVariant : uint32_t { TileLds = 0, QuadWave = 1, QuadWaveXcd = 2 };
void launch_raster(RasterVariant variant, const float4 *rec, const uint2 *rect, const uint2 *ranges,
                   const uint32_t *list, uint32_t width, uint32_t height, uint32_t tiles_x, uint32_t tiles_y,
                   float4 *out32, uchar4 *out8, hipStream_t stream);

// Device fill of n u32 words (graph-safe replacement of hipMemsetAsync).
void launch_fill_u32(uint32_t *p, uint32_t value, size_t n, hipStream_t stream);

}  // namespace gs
