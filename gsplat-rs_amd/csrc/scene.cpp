// scene.cpp -- camera math, scene validation and load-time preparation.
//
// Everything here runs once per scene load or per camera change on the host
// and is compiled with contraction off, so the values handed to the kernels
// (view rows, alpha, world covariance) equal the reference's IEEE
// single-precision results.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>

#include "common.h"

namespace gs {
namespace {

struct Quat {
  float x, y, z, w;
};

// quat_normalize (render-wgpu/src/lib.rs:1380-1387).
Quat normalize(Quat q) {
  float n2 = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
  if (n2 <= 0.0f) return Quat{0.0f, 0.0f, 0.0f, 1.0f};
  float inv = 1.0f / std::sqrt(n2);
  return Quat{q.x * inv, q.y * inv, q.z * inv, q.w * inv};
}

// quat_to_mat3 (lib.rs:1389-1407); normalises its argument first.
void to_mat3(Quat q, float m[9]) {
  q = normalize(q);
  const float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
  const float xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
  const float wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
  m[0] = 1.0f - 2.0f * (yy + zz);
  m[1] = 2.0f * (xy - wz);
  m[2] = 2.0f * (xz + wy);
  m[3] = 2.0f * (xy + wz);
  m[4] = 1.0f - 2.0f * (xx + zz);
  m[5] = 2.0f * (yz - wx);
  m[6] = 2.0f * (xz - wy);
  m[7] = 2.0f * (yz + wx);
  m[8] = 1.0f - 2.0f * (xx + yy);
}

// mat3_mul (lib.rs:1409-1417): row times column, every product summed.
void mul3(const float a[9], const float b[9], float o[9]) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) o[3 * r + c] = a[3 * r] * b[c] + a[3 * r + 1] * b[3 + c] + a[3 * r + 2] * b[6 + c];
}

bool finite_all(const std::vector<float> &v) {
  for (float f : v)
    if (!std::isfinite(f)) return false;
  return true;
}

struct V3 {
  float x, y, z;
};
V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
// vec3_normalize (ffi-c/src/lib.rs:1761-1768).
bool normalize3(V3 v, V3 *out) {
  float len = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
  if (!std::isfinite(len) || len <= 1.0e-6f) return false;
  float s = 1.0f / len;
  *out = V3{v.x * s, v.y * s, v.z * s};
  return true;
}

}  // namespace

GsplatCamera camera_default() {
  GsplatCamera c;
  c.position[0] = c.position[1] = c.position[2] = 0.0f;
  c.rotation_xyzw[0] = c.rotation_xyzw[1] = c.rotation_xyzw[2] = 0.0f;
  c.rotation_xyzw[3] = 1.0f;
  // 60_f32.to_radians() (core/src/lib.rs:67): 60 * (PI/180) in f32.
  c.vertical_fov_radians = 60.0f * (3.14159265358979323846f / 180.0f);
  c.near_plane = 0.01f;
  c.far_plane = 1000.0f;
  return c;
}

bool camera_valid(const GsplatCamera &c) {
  for (float v : c.position)
    if (!std::isfinite(v)) return false;
  for (float v : c.rotation_xyzw)
    if (!std::isfinite(v)) return false;
  float n2 = 0.0f;
  for (float v : c.rotation_xyzw) n2 += v * v;
  if (n2 <= 0.0f || !std::isfinite(n2)) return false;
  if (std::fabs(n2 - 1.0f) > 1.0e-3f) return false;
  if (!std::isfinite(c.vertical_fov_radians) || !std::isfinite(c.near_plane) || !std::isfinite(c.far_plane))
    return false;
  if (c.vertical_fov_radians <= 0.0f || c.vertical_fov_radians >= 3.14159265358979323846f ||
      c.near_plane <= 0.0f || c.far_plane <= c.near_plane)
    return false;
  return true;
}

void view_rows(const float q[4], float rows[9]) {
  Quat n = normalize(Quat{q[0], q[1], q[2], q[3]});  // quat_inverse (lib.rs:1374-1378)
  to_mat3(Quat{-n.x, -n.y, -n.z, n.w}, rows);
}

void rotation_looking_at(const float position[3], const float target[3], float out[4]) {
  V3 p{position[0], position[1], position[2]}, t{target[0], target[1], target[2]};
  V3 forward;
  if (!normalize3(sub(t, p), &forward)) forward = V3{0.0f, 0.0f, 1.0f};
  V3 world_up = std::fabs(forward.y) > 0.98f ? V3{0.0f, 0.0f, 1.0f} : V3{0.0f, 1.0f, 0.0f};
  V3 right;
  if (!normalize3(cross(world_up, forward), &right)) right = V3{1.0f, 0.0f, 0.0f};
  V3 up = cross(forward, right);
  // quat_from_camera_basis (ffi-c/src/lib.rs:1710-1739).
  float m00 = right.x, m01 = up.x, m02 = forward.x;
  float m10 = right.y, m11 = up.y, m12 = forward.y;
  float m20 = right.z, m21 = up.z, m22 = forward.z;
  float trace = m00 + m11 + m22;
  float x, y, z, w;
  if (trace > 0.0f) {
    float s = std::sqrt(trace + 1.0f) * 2.0f;
    x = (m21 - m12) / s;
    y = (m02 - m20) / s;
    z = (m10 - m01) / s;
    w = 0.25f * s;
  } else if (m00 > m11 && m00 > m22) {
    float s = std::sqrt(1.0f + m00 - m11 - m22) * 2.0f;
    x = 0.25f * s;
    y = (m01 + m10) / s;
    z = (m02 + m20) / s;
    w = (m21 - m12) / s;
  } else if (m11 > m22) {
    float s = std::sqrt(1.0f + m11 - m00 - m22) * 2.0f;
    x = (m01 + m10) / s;
    y = 0.25f * s;
    z = (m12 + m21) / s;
    w = (m02 - m20) / s;
  } else {
    float s = std::sqrt(1.0f + m22 - m00 - m11) * 2.0f;
    x = (m02 + m20) / s;
    y = (m12 + m21) / s;
    z = 0.25f * s;
    w = (m10 - m01) / s;
  }
  float norm = std::max(std::sqrt(x * x + y * y + z * z + w * w), 1.0e-6f);
  out[0] = x / norm;
  out[1] = y / norm;
  out[2] = z / norm;
  out[3] = w / norm;
}

// auto_surface_camera_control + surface_camera_from_control with yaw = pitch
// = 0 (ffi-c/src/lib.rs:1593-1675).
GsplatCamera auto_camera(const HostScene *scene, uint32_t width, uint32_t height) {
  if (scene == nullptr) throw Error(Code::SceneNotLoaded, "scene is not loaded");
  if (scene->count == 0) throw Error(Code::InvalidArgument, "scene is not loaded");
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = 0; i < scene->count; ++i)
    for (int a = 0; a < 3; ++a) {
      float v = scene->positions[3 * (size_t)i + a];
      mn[a] = std::fmin(mn[a], v);  // f32::min ignores NaN like fmin
      mx[a] = std::fmax(mx[a], v);
    }
  float center[3] = {(mn[0] + mx[0]) * 0.5f, (mn[1] + mx[1]) * 0.5f, (mn[2] + mx[2]) * 0.5f};
  float half_x = std::max((mx[0] - mn[0]) * 0.5f, 1e-3f);
  float half_y = std::max((mx[1] - mn[1]) * 0.5f, 1e-3f);
  float half_z = std::max((mx[2] - mn[2]) * 0.5f, 1e-3f);
  float aspect = (float)width / (float)height;
  float vfov = std::max(camera_default().vertical_fov_radians, 1e-3f);
  float hfov = 2.0f * std::atan(std::tan(vfov * 0.5f) * aspect);
  float dist_y = half_y / std::tan(vfov * 0.5f);
  float dist_x = half_x / std::tan(hfov * 0.5f);
  float dist = (std::max(dist_y, dist_x) + half_z) * 1.2f;
  float radius = std::max(std::max(half_x, half_y), half_z);

  GsplatCamera cam = camera_default();
  // pitch = 0, yaw = 0: offset = (sin 0 * cos 0 * d, sin 0 * d, -cos 0 * cos 0 * d).
  const float pitch = 0.0f, yaw = 0.0f;
  float cos_pitch = std::cos(pitch);
  float offset[3] = {std::sin(yaw) * cos_pitch * dist, std::sin(pitch) * dist, -std::cos(yaw) * cos_pitch * dist};
  for (int a = 0; a < 3; ++a) cam.position[a] = center[a] + offset[a];
  rotation_looking_at(cam.position, center, cam.rotation_xyzw);
  float r = std::max(radius, 1.0e-3f);
  cam.near_plane = std::max(dist - r * 2.0f, 0.01f);
  cam.far_plane = std::max(dist + r * 8.0f, 100.0f);
  float a2 = std::max((float)width / (float)std::max(height, 1u), 1.0e-3f);
  if (a2 < 0.6f) cam.vertical_fov_radians = 65.0f * (3.14159265358979323846f / 180.0f);
  return cam;
}

void validate_scene(const HostScene &s) {
  size_t n = s.count;
  if (s.positions.size() != 3 * n || s.opacity.size() != n || s.scale_xyz.size() != 3 * n ||
      s.rotation_xyzw.size() != 4 * n || s.color_dc.size() != 3 * n)
    throw Error(Code::ParseFailed, "invalid scene buffers");
  if (!finite_all(s.positions) || !finite_all(s.opacity) || !finite_all(s.scale_xyz) ||
      !finite_all(s.rotation_xyzw) || !finite_all(s.color_dc))
    throw Error(Code::ParseFailed, "invalid scene buffers");
  if (s.sh_degree > 4) throw Error(Code::Unsupported, "invalid scene buffers");
  size_t stride = 3 * (size_t)sh_per_channel(s.sh_degree);
  if (!s.has_sh_rest) {
    if (stride != 0) throw Error(Code::ParseFailed, "invalid scene buffers");
  } else {
    if (stride == 0 || s.sh_rest.size() != n * stride || !finite_all(s.sh_rest))
      throw Error(Code::ParseFailed, "invalid scene buffers");
  }
}

HostScene scene_from_view(const GsplatSceneView &v) {
  HostScene s;
  s.count = v.count;
  s.sh_degree = v.sh_degree;
  size_t n = v.count;
  if (n > 0 && (!v.positions || !v.opacity || !v.scale_xyz || !v.rotation_xyzw || !v.color_dc))
    throw Error(Code::InvalidArgument, "scene view has a null buffer");
  if (v.sh_degree > 4) throw Error(Code::Unsupported, "invalid scene buffers");
  auto cp = [n](const float *p, size_t k) { return p ? std::vector<float>(p, p + n * k) : std::vector<float>(); };
  s.positions = cp(v.positions, 3);
  s.opacity = cp(v.opacity, 1);
  s.scale_xyz = cp(v.scale_xyz, 3);
  s.rotation_xyzw = cp(v.rotation_xyzw, 4);
  s.color_dc = cp(v.color_dc, 3);
  size_t stride = 3 * (size_t)sh_per_channel(v.sh_degree);
  s.has_sh_rest = v.sh_rest != nullptr;
  if (s.has_sh_rest) s.sh_rest.assign(v.sh_rest, v.sh_rest + n * stride);
  validate_scene(s);
  return s;
}

float coverage_cut(float alpha) {
  auto pass = [alpha](float r2) { return r2 <= 1.0f && alpha * std::exp(-4.5f * r2) > 1.0f / 256.0f; };
  if (!pass(0.0f)) return -1.0f;
  if (pass(1.0f)) return 1.0f;
  // The predicate is monotone in r2, so bisect on the IEEE bit pattern of the
  // non-negative float r2 for the last passing value.
  uint32_t lo = 0, hi = 0x3F800000u;
  while (hi - lo > 1) {
    uint32_t mid = lo + (hi - lo) / 2;
    float f;
    std::memcpy(&f, &mid, 4);
    if (pass(f))
      lo = mid;
    else
      hi = mid;
  }
  float out;
  std::memcpy(&out, &lo, 4);
  return out;
}

namespace {

// Round half away from zero (Rust's f32::round) of a non-negative value.
uint32_t round_u(float x) { return (uint32_t)std::round(x); }

// SceneBounds::encode_u16 and LogScaleRange::encode_u10 (packed_atlas.rs:52-190).
uint32_t quantize(float v, float lo, float span, float levels) {
  const float t = std::min(std::max((v - lo) / span, 0.0f), 1.0f);
  return round_u(t * levels);
}

// pack_quat_smallest_three (packed_atlas.rs:273-309): the largest component
// is dropped (made positive), the other three map [-1/sqrt 2, 1/sqrt 2] to
// 10 bits each above a 2-bit index.
uint32_t pack_quat(const float *xyzw) {
  float q[4] = {xyzw[0], xyzw[1], xyzw[2], xyzw[3]};
  const float len2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  if (len2 <= 1e-20f) {
    q[0] = q[1] = q[2] = 0.0f;
    q[3] = 1.0f;
  } else {
    const float inv = 1.0f / std::sqrt(len2);
    for (float &c : q) c *= inv;
  }
  uint32_t big = 0;
  for (uint32_t k = 1; k < 4; ++k)
    if (std::fabs(q[k]) > std::fabs(q[big])) big = k;
  if (q[big] < 0.0f)
    for (float &c : q) c = -c;
  const float scale = 0.70710678118654752440f;
  uint32_t out = big, lane = 0;
  for (uint32_t k = 0; k < 4; ++k) {
    if (k == big) continue;
    const float t = std::min(std::max((q[k] / scale) * 0.5f + 0.5f, 0.0f), 1.0f);
    out |= (round_u(t * 1023.0f) & 0x3FFu) << (2 + 10 * lane++);
  }
  return out;
}

}  // namespace

PackedScene pack_scene(const HostScene &s) {
  PackedScene p;
  p.count = s.count;
  p.sh_degree = s.has_sh_rest ? s.sh_degree : 0;
  if (p.sh_degree > 3) throw Error(Code::Unsupported, "the packed scene codec holds SH degrees 0..3");
  const size_t n = s.count;
  const uint32_t pc = p.sh_degree ? sh_per_channel(p.sh_degree) : 0;
  p.consts.assign(kPackedConsts, 0.0f);
  float *k = p.consts.data();
  // SceneBounds::from_positions, LogScaleRange::from_scales, scene_sh_scales
  // (packed_atlas.rs:52-78, 160-190, 439-462).
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  float lmin = INFINITY, lmax = -INFINITY;
  float shs[3] = {1e-6f, 1e-6f, 1e-6f};
  for (size_t i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], s.positions[3 * i + a]);
      hi[a] = std::max(hi[a], s.positions[3 * i + a]);
      lmin = std::min(lmin, s.scale_xyz[3 * i + a]);
      lmax = std::max(lmax, s.scale_xyz[3 * i + a]);
    }
    for (uint32_t ch = 0; ch < 3; ++ch)
      for (uint32_t c = 0; c < pc; ++c) shs[ch] = std::max(shs[ch], std::fabs(s.sh_rest[i * 3 * pc + ch * pc + c]));
  }
  for (int a = 0; a < 3; ++a) {
    if (!std::isfinite(lo[a]) || !std::isfinite(hi[a]) || std::fabs(hi[a] - lo[a]) < 1e-6f) {
      lo[a] = 0.0f;
      hi[a] = 1.0f;
    }
    k[kPkMin + a] = lo[a];
    k[kPkSpan + a] = std::max(hi[a] - lo[a], 1e-6f);
  }
  if (!std::isfinite(lmin) || !std::isfinite(lmax)) {
    lmin = -8.0f;
    lmax = 8.0f;
  } else if (std::fabs(lmax - lmin) < 1e-6f) {
    lmax = lmin + 1e-3f;
  }
  k[kPkLogMin] = lmin;
  k[kPkLogSpan] = std::max(lmax - lmin, 1e-6f);
  for (int ch = 0; ch < 3; ++ch) k[kPkShScale + ch] = shs[ch];
  // decode_opacity of the packed shader (splat_surface_packed.wgsl:130-133):
  // alpha = clamp(code / 255, 1e-4, 1 - 1e-4); its coverage cut per code.
  for (uint32_t c = 0; c < 256; ++c)
    k[kPkCut + c] = coverage_cut(std::min(std::max((float)c / 255.0f, 1e-4f), 1.0f - 1e-4f));

  const uint32_t words = packed_words(p.sh_degree);
  p.words.assign((size_t)words * n, 0u);
  uint32_t *w = p.words.data();
  const float C0 = 0.28209479177387814f;
  for (size_t i = 0; i < n; ++i) {
    uint32_t pos[3];
    for (int a = 0; a < 3; ++a) pos[a] = quantize(s.positions[3 * i + a], k[kPkMin + a], k[kPkSpan + a], 65535.0f);
    // encode_opacity_u8 (packed_atlas.rs:268-271)
    const float alpha = std::min(std::max(1.0f / (1.0f + std::exp(-s.opacity[i])), 0.0f), 1.0f);
    const uint32_t op = round_u(alpha * 255.0f);
    uint32_t scales = 0;
    for (int a = 0; a < 3; ++a)
      scales |= quantize(s.scale_xyz[3 * i + a], lmin, k[kPkLogSpan], 1023.0f) << (10 * a);
    uint32_t rgb = 0;  // pack_color_rgb10 of the degree-0 bake (packed_atlas.rs:311-318, 517-524)
    for (int a = 0; a < 3; ++a) {
      const float v = std::min(std::max(C0 * s.color_dc[3 * i + a] + 0.5f, 0.0f), 1.0f);
      rgb |= (round_u(v * 1023.0f) & 0x3FFu) << (10 * a);
    }
    w[0 * n + i] = pos[0] | (pos[1] << 16);
    w[1 * n + i] = pos[2] | (op << 16);
    w[2 * n + i] = scales;
    w[3 * n + i] = pack_quat(&s.rotation_xyzw[4 * i]);
    w[4 * n + i] = rgb;
    if (pc) {  // quantize_sh_rest_with_scales (packed_atlas.rs:380-400): 15 lanes per channel
      uint8_t b[48] = {};
      for (uint32_t ch = 0; ch < 3; ++ch)
        for (uint32_t c = 0; c < pc; ++c) {
          const float v = s.sh_rest[i * 3 * pc + ch * pc + c] / shs[ch] * 127.0f;
          const float r = std::min(std::max(std::round(v), -127.0f), 127.0f);
          b[ch * 15 + c] = (uint8_t)(int8_t)r;
        }
      for (uint32_t q = 0; q < kPackedShWords; ++q) {
        uint32_t word;
        std::memcpy(&word, b + 4 * q, 4);
        w[(kPackedHotWords + q) * n + i] = word;
      }
    }
  }
  return p;
}

// precompute_alpha_values (lib.rs:1251-1261), precompute_world_covariances
// (lib.rs:1654-1674), and the coverage cut, into the device SoA image.
PreparedScene prepare_scene(const HostScene &s) {
  PreparedScene p;
  p.count = s.count;
  p.sh_degree = s.has_sh_rest ? s.sh_degree : 0;
  const size_t n = s.count;
  const uint32_t pc = p.sh_degree ? sh_per_channel(p.sh_degree) : 0;
  const uint32_t fields = soa_fields(p.sh_degree);
  p.soa.assign((size_t)fields * n, 0.0f);
  float *soa = p.soa.data();
  auto field = [soa, n](uint32_t f) { return soa + (size_t)f * n; };

  auto work = [&](size_t begin, size_t end) {
    for (size_t i = begin; i < end; ++i) {
      for (int a = 0; a < 3; ++a) field(kPx + a)[i] = s.positions[3 * i + a];
      float sig = 1.0f / (1.0f + std::exp(-s.opacity[i]));
      float alpha = std::min(std::max(sig, 0.0f), 1.0f);
      field(kAlpha)[i] = alpha;
      field(kCut)[i] = coverage_cut(alpha);

      float sx = std::max(std::exp(s.scale_xyz[3 * i + 0]), 1e-6f);
      float sy = std::max(std::exp(s.scale_xyz[3 * i + 1]), 1e-6f);
      float sz = std::max(std::exp(s.scale_xyz[3 * i + 2]), 1e-6f);
      float obj[9] = {sx * sx, 0.0f, 0.0f, 0.0f, sy * sy, 0.0f, 0.0f, 0.0f, sz * sz};
      const float *q = &s.rotation_xyzw[4 * i];
      float rot[9], rt[9], tmp[9], w[9];
      to_mat3(normalize(Quat{q[0], q[1], q[2], q[3]}), rot);
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) rt[3 * r + c] = rot[3 * c + r];
      mul3(rot, obj, tmp);
      mul3(tmp, rt, w);
      field(kCxx)[i] = w[0];
      field(kCxy)[i] = w[1];
      field(kCxz)[i] = w[2];
      field(kCyy)[i] = w[4];
      field(kCyz)[i] = w[5];
      field(kCzz)[i] = w[8];
      for (int a = 0; a < 3; ++a) field(kDcR + a)[i] = s.color_dc[3 * i + a];
      for (uint32_t ch = 0; ch < 3; ++ch)
        for (uint32_t k = 0; k < pc; ++k) field(kRest0 + ch * pc + k)[i] = s.sh_rest[i * 3 * pc + ch * pc + k];
    }
  };
  unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 65536 || hw == 1) {
    work(0, n);
  } else {
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < hw; ++t) pool.emplace_back(work, n * t / hw, n * (t + 1) / hw);
    for (auto &th : pool) th.join();
  }
  return p;
}

}  // namespace gs
