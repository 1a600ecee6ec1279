// common.h -- shared host-side types of the MI355X SortedAlpha renderer.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "gsplat.h"
#include "gsplat_ext.h"

namespace gs {

// ErrorCode of the reference (crates/gsplat-core/src/lib.rs:154-164).
enum class Code : int32_t {
  Ok = 0,
  InvalidArgument = 1,
  NotFound = 2,
  ParseFailed = 3,
  Unsupported = 4,
  SceneNotLoaded = 5,
  Internal = 100,
};

// Error carried to the C ABI boundary; `what()` is the detail text appended
// after "<function>: " in the thread-local last-error message.
struct Error : std::runtime_error {
  Code code;
  Error(Code c, const std::string &msg) : std::runtime_error(msg), code(c) {}
};

// Host copy of a scene in the reference's SceneBuffers layout
// (core/src/lib.rs:201-213).
struct HostScene {
  uint32_t count = 0;
  uint32_t sh_degree = 0;
  std::vector<float> positions;      // 3N
  std::vector<float> opacity;        // N
  std::vector<float> scale_xyz;      // 3N
  std::vector<float> rotation_xyzw;  // 4N
  std::vector<float> color_dc;       // 3N
  std::vector<float> sh_rest;        // N*3*per_channel, empty when degree 0
  bool has_sh_rest = false;
};

inline uint32_t sh_per_channel(uint32_t degree) { return (degree + 1) * (degree + 1) - 1; }

// Camera of gsplat.h with the reference's default values.
GsplatCamera camera_default();
// Camera::validate (core/src/lib.rs:39-107).
bool camera_valid(const GsplatCamera &c);
// quat_to_mat3(quat_inverse(q)) rows (render-wgpu/src/lib.rs:1374-1407).
void view_rows(const float q_xyzw[4], float rows[9]);
// Auto-framing camera of gsplat_context_set_auto_camera
// (ffi-c/src/lib.rs:1593-1675); throws SceneNotLoaded / InvalidArgument.
GsplatCamera auto_camera(const HostScene *scene, uint32_t width, uint32_t height);
// camera_rotation_looking_at (ffi-c/src/lib.rs:1677-1733).
void rotation_looking_at(const float position[3], const float target[3], float out_xyzw[4]);

// SceneBuffers::validate (core/src/lib.rs:224-276).
void validate_scene(const HostScene &s);
// Copies a borrowed view into a HostScene (validating it).
HostScene scene_from_view(const GsplatSceneView &v);

// Per-splat load-time values (render-wgpu/src/lib.rs:1251-1261, 1654-1674)
// plus the coverage cut of the raster test.
struct PreparedScene {
  uint32_t count = 0;
  uint32_t sh_degree = 0;
  std::vector<float> soa;  // device image, see scene_layout()
};

// Device SoA layout of a prepared scene: float arrays of length N each,
// in this order: px py pz | cxx cxy cxz cyy cyz czz | alpha cut | dc_r dc_g
// dc_b | rest[ch*per_channel + k] for ch in 0..3, k in 0..per_channel.
enum SoaField : uint32_t {
  kPx = 0, kPy, kPz,
  kCxx, kCxy, kCxz, kCyy, kCyz, kCzz,
  kAlpha, kCut,
  kDcR, kDcG, kDcB,
  kRest0,
};
inline uint32_t soa_fields(uint32_t degree) { return kRest0 + 3 * sh_per_channel(degree); }

PreparedScene prepare_scene(const HostScene &s);

// Packed scene codec (the reference's packed atlas, render-wgpu/src/
// packed_atlas.rs:1-600, shaders/splat_surface_packed.wgsl:125-241): per splat
// a 20-byte hot record -- u16 bounds-relative position, u8 opacity (alpha),
// three u10 log-scales over the scene's range, a smallest-three quaternion
// (10+10+10+2 bits), the DC colour baked to RGB10 -- and, for SH degrees 1..3,
// a 48-byte sidecar of 45 i8 rest coefficients over per-channel scene scales.
// Stored coefficient-major: word w of splat i at words[w * N + i].
constexpr uint32_t kPackedHotWords = 5;
constexpr uint32_t kPackedShWords = 12;
constexpr uint32_t packed_words(uint32_t degree) { return kPackedHotWords + (degree ? kPackedShWords : 0); }
// Scene-wide decode constants (device array of kPackedConsts floats):
// bounds min xyz, bounds span xyz, log-scale min, log-scale span, SH scales
// rgb, then the coverage cut of each of the 256 opacity codes.
enum PackedConst : uint32_t { kPkMin = 0, kPkSpan = 3, kPkLogMin = 6, kPkLogSpan = 7, kPkShScale = 8, kPkCut = 16 };
constexpr uint32_t kPackedConsts = kPkCut + 256;
struct PackedScene {
  uint32_t count = 0;
  uint32_t sh_degree = 0;
  std::vector<uint32_t> words;
  std::vector<float> consts;  // kPackedConsts
};
// Throws Unsupported for SH degree 4 (the sidecar holds degrees up to 3).
PackedScene pack_scene(const HostScene &s);

// Largest r2 in [0,1] for which a splat of this alpha passes the fragment
// test r2 <= 1 && alpha*exp(-4.5 r2) > 1/256 (splat_surface_direct.wgsl:358-367),
// or -1 when none does.
float coverage_cut(float alpha);

// PLY loader (crates/gsplat-io-ply).
struct PlyLimits {
  uint64_t max_input_bytes = 1ull << 30;
  uint64_t max_header_bytes = 1ull << 20;
  uint64_t max_vertices = 5000000;
  uint64_t max_vertex_properties = 128;
  uint64_t max_scene_bytes = 1ull << 30;
};
HostScene ply_load(const std::string &path, const PlyLimits &limits);
HostScene ply_parse(const uint8_t *bytes, size_t len, const PlyLimits &limits, bool text_only);

}  // namespace gs
