// kernels.hip -- gfx950 kernels of the SortedAlpha path, over a batch of views.
//
// Compiled with -ffp-contract=off and correctly rounded f32 divide/sqrt, so
// every expression that restates reference arithmetic (depth key,
// projection, SH colour, pixel bounds) rounds exactly like the reference's
// IEEE single precision; fused multiply-adds appear only where written as
// fmaf (the per-pixel r^2 of the raster, whose operation order is defined by
// this renderer because the reference leaves it to the GPU's interpolator).
//
// A view's splats occupy a slot segment aligned to kSegmentAlign, so every
// 256-thread workgroup over slots belongs to exactly one view, and the view
// is found once per workgroup by binary search over the segment starts.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <type_traits>

#include "common.h"
#include "kernels.h"
#include "sort.h"

namespace gs {

#define GS_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw Error(Code::Internal, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace {

constexpr uint32_t kWave = 64;
constexpr uint32_t kWaves = kThreads / kWave;

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
  const uint32_t lane = __lane_id();
#pragma unroll
  for (uint32_t d = 1; d < kWave; d <<= 1) {
    uint32_t t = __shfl_up(v, d, kWave);
    if (lane >= d) v += t;
  }
  return v;
}

// Exclusive scan across the 256 threads of a workgroup; *total = block sum.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *lds4, uint32_t *total) {
  const uint32_t inc = wave_inclusive_scan(v);
  const uint32_t w = threadIdx.x / kWave;
  if (__lane_id() == kWave - 1) lds4[w] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (uint32_t i = 0; i < kWaves; ++i) {
    uint32_t x = lds4[i];
    off += (i < w) ? x : 0u;
    tot += x;
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

// Index v of the last start[v] <= x (start is non-decreasing, start[0] = 0).
__device__ __forceinline__ uint32_t find_view(const uint32_t *__restrict__ start, uint32_t nviews, uint32_t x) {
  uint32_t lo = 0, hi = nviews - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) / 2;
    if (start[mid] <= x)
      lo = mid;
    else
      hi = mid - 1;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

// ---------------------------------------------------------------- projection

struct Cov6 {
  float xx, xy, xz, yy, yz, zz;
};

// covariance_quadratic / covariance_bilinear (splat_surface_direct.wgsl:95-109).
__device__ __forceinline__ float cov_quadratic(const Cov6 &c, float r0, float r1, float r2) {
  return r0 * r0 * c.xx + 2.0f * r0 * r1 * c.xy + 2.0f * r0 * r2 * c.xz + r1 * r1 * c.yy +
         2.0f * r1 * r2 * c.yz + r2 * r2 * c.zz;
}
__device__ __forceinline__ float cov_bilinear(const Cov6 &c, float a0, float a1, float a2, float b0, float b1,
                                              float b2) {
  float bx = c.xx * b0 + c.xy * b1 + c.xz * b2;
  float by = c.xy * b0 + c.yy * b1 + c.yz * b2;
  float bz = c.xz * b0 + c.yz * b1 + c.zz * b2;
  return a0 * bx + a1 * by + a2 * bz;
}

constexpr float kShC0 = 0.28209479177387814f;

// One splat of eval_sh_rgb (wgsl:233-326) for channel ch, in the shader's
// left-to-right accumulation order, from the degree-0 term C0 * dc
// (src.dc_term) and the channel's coefficient k = src.rest(base + k).
template <int DEG, class Src>
__device__ __forceinline__ float sh_channel(const Src &src, uint32_t ch, float x, float y, float z) {
  constexpr uint32_t PC = (DEG + 1) * (DEG + 1) - 1;
  const uint32_t base = ch * PC;
  float r = src.dc_term(ch);
  if constexpr (DEG >= 1) {
    const float C1 = 0.4886025119029199f;
    r = r - C1 * y * src.rest(base + 0) + C1 * z * src.rest(base + 1) - C1 * x * src.rest(base + 2);
  }
  if constexpr (DEG >= 2) {
    const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
    r = r + 1.0925484305920792f * xy * src.rest(base + 3) + -1.0925484305920792f * yz * src.rest(base + 4) +
        0.31539156525252005f * (2.0f * zz - xx - yy) * src.rest(base + 5) +
        -1.0925484305920792f * xz * src.rest(base + 6) + 0.5462742152960396f * (xx - yy) * src.rest(base + 7);
    if constexpr (DEG >= 3) {
      r = r + -0.5900435899266435f * y * (3.0f * xx - yy) * src.rest(base + 8) +
          2.890611442640554f * xy * z * src.rest(base + 9) +
          -0.4570457994644658f * y * (4.0f * zz - xx - yy) * src.rest(base + 10) +
          0.3731763325901154f * z * (2.0f * zz - 3.0f * xx - 3.0f * yy) * src.rest(base + 11) +
          -0.4570457994644658f * x * (4.0f * zz - xx - yy) * src.rest(base + 12) +
          1.445305721320277f * z * (xx - yy) * src.rest(base + 13) +
          -0.5900435899266435f * x * (xx - 3.0f * yy) * src.rest(base + 14);
    }
    if constexpr (DEG >= 4) {
      r = r + 2.5033429417967046f * xy * (xx - yy) * src.rest(base + 15) +
          -1.7701307697799304f * yz * (3.0f * xx - yy) * src.rest(base + 16) +
          0.9461746957575601f * xy * (7.0f * zz - 1.0f) * src.rest(base + 17) +
          -0.6690465435572892f * yz * (7.0f * zz - 3.0f) * src.rest(base + 18) +
          0.10578554691520431f * (zz * (35.0f * zz - 30.0f) + 3.0f) * src.rest(base + 19) +
          -0.6690465435572892f * xz * (7.0f * zz - 3.0f) * src.rest(base + 20) +
          0.47308734787878004f * (xx - yy) * (7.0f * zz - 1.0f) * src.rest(base + 21) +
          -1.7701307697799304f * xz * (xx - 3.0f * yy) * src.rest(base + 22) +
          0.6258357354491761f * (xx * (xx - 3.0f * yy) - yy * (3.0f * xx - yy)) * src.rest(base + 23);
    }
  }
  return r;
}

// A splat's scene quantities as the projection reads them: position,
// world covariance, alpha, coverage cut, the degree-0 colour term and the
// SH rest coefficients (channel-major). Three sources provide them.

// Fields of splat i of an fp32 SoA scene (common.h order), read where the
// projection needs them (culled splats skip their covariance and SH loads).
template <int DEG>
struct SoaSource {
  const float *__restrict__ soa;
  size_t n;
  uint32_t i;
  __device__ __forceinline__ float f(uint32_t k) const { return soa[k * n + i]; }
  __device__ __forceinline__ float pos(uint32_t a) const { return f(kPx + a); }
  __device__ __forceinline__ Cov6 cov() const { return {f(kCxx), f(kCxy), f(kCxz), f(kCyy), f(kCyz), f(kCzz)}; }
  __device__ __forceinline__ float alpha() const { return f(kAlpha); }
  __device__ __forceinline__ float cut() const { return f(kCut); }
  __device__ __forceinline__ float dc_term(uint32_t ch) const { return kShC0 * f(kDcR + ch); }
  __device__ __forceinline__ float rest(uint32_t k) const { return f(kRest0 + k); }
};

// Decode of the packed scene codec (common.h pack_scene), following the
// packed shader (splat_surface_packed.wgsl:125-241) and the codec's CPU
// decoders (packed_atlas.rs:86-97, 184-189, 311-345, 409-419). Word q of splat
// i is w[q * n + i]; pk holds the scene's decode constants. The codec is
// lossy, and its divisions by 65535, 1023, 255 and 127 are multiplications by
// the reciprocal here (within an ulp of the quotient, as GPU shading
// languages allow for `/`).
struct PackedWords {
  const uint32_t *__restrict__ w;
  const float *__restrict__ pk;
  size_t n;
  uint32_t i;
  __device__ __forceinline__ uint32_t word(uint32_t q) const { return w[q * n + i]; }
  // decode_position: min + (code / 65535) * extent.
  __device__ __forceinline__ float pos(uint32_t a) const {
    const uint32_t code = a == 0 ? (word(0) & 0xFFFFu) : a == 1 ? (word(0) >> 16) : (word(1) & 0xFFFFu);
    return pk[kPkMin + a] + ((float)code * (1.0f / 65535.0f)) * pk[kPkSpan + a];
  }
  __device__ __forceinline__ uint32_t opacity_code() const { return (word(1) >> 16) & 0xFFu; }
  // decode_opacity: clamp(code / 255, 1e-4, 1 - 1e-4).
  __device__ __forceinline__ float alpha() const {
    return fminf(fmaxf((float)opacity_code() * (1.0f / 255.0f), 1e-4f), 1.0f - 1e-4f);
  }
  __device__ __forceinline__ float cut() const { return pk[kPkCut + opacity_code()]; }
  // world_covariance_from_scale_rotation of decode_log_scale and
  // unpack_quat_smallest_three, with quat_to_mat3's rows.
  __device__ __forceinline__ Cov6 cov() const {
    const uint32_t sw = word(2), qw = word(3);
    float s2[3];
#pragma unroll
    for (uint32_t a = 0; a < 3; ++a) {
      const float ls = pk[kPkLogMin] + ((float)((sw >> (10 * a)) & 0x3FFu) * (1.0f / 1023.0f)) * pk[kPkLogSpan];
      s2[a] = expf(2.0f * ls);
    }
    const float inv_sqrt2 = 0.70710678118654752440f;
    float c[3];
#pragma unroll
    for (uint32_t l = 0; l < 3; ++l)
      c[l] = ((float)((qw >> (2 + 10 * l)) & 0x3FFu) * (1.0f / 1023.0f) * 2.0f - 1.0f) * inv_sqrt2;
    const float dropped = sqrtf(fmaxf(1.0f - (c[0] * c[0] + c[1] * c[1] + c[2] * c[2]), 0.0f));
    const uint32_t big = qw & 3u;
    float q[4];
    q[0] = big == 0 ? dropped : c[0];
    q[1] = big == 0 ? c[0] : big == 1 ? dropped : c[1];
    q[2] = big <= 1 ? c[1] : big == 2 ? dropped : c[2];
    q[3] = big == 3 ? dropped : c[2];
    const float len2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    float x = 0.0f, y = 0.0f, z = 0.0f, ww = 1.0f;
    if (len2 > 1e-20f) {
      const float inv = 1.0f / sqrtf(len2);
      x = q[0] * inv;
      y = q[1] * inv;
      z = q[2] * inv;
      ww = q[3] * inv;
    }
    const float x2 = x + x, y2 = y + y, z2 = z + z;
    const float xx = x * x2, yy = y * y2, zz = z * z2, xy = x * y2, xz = x * z2, yz = y * z2;
    const float wx = ww * x2, wy = ww * y2, wz = ww * z2;
    const float r0[3] = {1.0f - (yy + zz), xy - wz, xz + wy};
    const float r1[3] = {xy + wz, 1.0f - (xx + zz), yz - wx};
    const float r2[3] = {xz - wy, yz + wx, 1.0f - (xx + yy)};
    auto term = [&](const float *a, const float *b) {
      return s2[0] * a[0] * b[0] + s2[1] * a[1] * b[1] + s2[2] * a[2] * b[2];
    };
    return {term(r0, r0), term(r0, r1), term(r0, r2), term(r1, r1), term(r1, r2), term(r2, r2)};
  }
  // The RGB10 degree-0 bake, clamp(C0 * dc + 0.5, 0, 1), less the 0.5 the
  // colour adds back.
  __device__ __forceinline__ float dc_term(uint32_t ch) const {
    return (float)((word(4) >> (10 * ch)) & 0x3FFu) * (1.0f / 1023.0f) - 0.5f;
  }
  // dequantize_sh_rest: byte ch * 15 + k of the 48-byte sidecar, q / 127 * scale.
  template <int DEG>
  __device__ __forceinline__ float rest_of(uint32_t k) const {
    constexpr uint32_t PC = (DEG + 1) * (DEG + 1) - 1;
    if constexpr (PC == 0) {
      return 0.0f;
    } else {
      const uint32_t ch = k / PC, b = ch * 15 + k % PC;
      const int8_t q = (int8_t)((word(kPackedHotWords + b / 4) >> (8 * (b % 4))) & 0xFFu);
      return (float)q * (pk[kPkShScale + ch] * (1.0f / 127.0f));
    }
  }
};

template <int DEG>
struct PackedSource : PackedWords {
  __device__ __forceinline__ float rest(uint32_t k) const { return this->template rest_of<DEG>(k); }
};

// All quantities of one splat in registers: read (and decoded) once,
// projected for every view of a group.
template <int DEG>
struct RegSource {
  static constexpr uint32_t kRest = 3 * ((DEG + 1) * (DEG + 1) - 1);
  float p[3], a, ct, dc[3], r[kRest > 0 ? kRest : 1];
  Cov6 c;
  template <class Src>
  __device__ __forceinline__ void load(const Src &s) {
#pragma unroll
    for (uint32_t k = 0; k < 3; ++k) p[k] = s.pos(k);
    c = s.cov();
    a = s.alpha();
    ct = s.cut();
#pragma unroll
    for (uint32_t k = 0; k < 3; ++k) dc[k] = s.dc_term(k);
#pragma unroll
    for (uint32_t k = 0; k < kRest; ++k) r[k] = s.rest(k);
  }
  __device__ __forceinline__ float pos(uint32_t k) const { return p[k]; }
  __device__ __forceinline__ Cov6 cov() const { return c; }
  __device__ __forceinline__ float alpha() const { return a; }
  __device__ __forceinline__ float cut() const { return ct; }
  __device__ __forceinline__ float dc_term(uint32_t ch) const { return dc[ch]; }
  __device__ __forceinline__ float rest(uint32_t k) const { return r[k]; }
};

// Scene format FMT (kernels.h scene_format) as a source of splat i of view
// or group `d` (anything with soa and n).
template <uint32_t FMT>
struct FormatOf {
  static constexpr bool kPacked = FMT >= kPackedFormat;
  static constexpr int kDeg = kPacked ? (int)(FMT - kPackedFormat) : (int)FMT;
  using Source = typename std::conditional<kPacked, PackedSource<kDeg>, SoaSource<kDeg>>::type;
  __device__ __forceinline__ static Source at(const float *soa, size_t n, uint32_t i) {
    if constexpr (kPacked) {
      const uint32_t *w = reinterpret_cast<const uint32_t *>(soa);
      Source s;
      s.w = w;
      s.pk = reinterpret_cast<const float *>(w + (size_t)packed_words(kDeg) * n);
      s.n = n;
      s.i = i;
      return s;
    } else {
      return Source{soa, n, i};
    }
  }
};

struct Projected {
  uint32_t key;  // depth key, 0 when culled by near/far
  bool drawn;    // project_splat did not return invalid_splat()
  float cx, cy, ux, uy, vx, vy, alpha;
  float col[3];  // premultiplied rgb
};

// Depth key and cull (preprocess_visible_into, render-wgpu/src/lib.rs:
// 1263-1296), projection (project_splat, wgsl:131-217) and colour
// (eval_color, wgsl:328-334) of splat i of a scene.
template <int DEG, class Src>
__device__ __forceinline__ Projected project_one(const FrameParams &fp, const Src &src) {
  Projected o{};
  const float relx = src.pos(0) - fp.cam[0];
  const float rely = src.pos(1) - fp.cam[1];
  const float relz = src.pos(2) - fp.cam[2];
  // world_to_camera_depth_with_view_row (lib.rs:1346-1354), no FMA.
  const float z = fp.rows[6] * relx + fp.rows[7] * rely + fp.rows[8] * relz;
  const bool vis = z >= fp.near_plane && z <= fp.far_plane;
  o.key = vis ? __float_as_uint(fmaxf(z, 0.0f)) : 0u;  // depth_to_key (lib.rs:973-976)
  if (!(vis && z > 1e-6f)) return o;
  const float x = fp.rows[0] * relx + fp.rows[1] * rely + fp.rows[2] * relz;
  const float y = fp.rows[3] * relx + fp.rows[4] * rely + fp.rows[5] * relz;
  const float inv_z = 1.0f / z;
  const float x_ndc = ((x * fp.f) * inv_z) / fp.aspect;
  const float y_ndc = (y * fp.f) * inv_z;
  const Cov6 wc = src.cov();
  const float *r0 = &fp.rows[0], *r1 = &fp.rows[3], *r2 = &fp.rows[6];
  // transform_covariance_terms_to_camera (wgsl:111-120).
  const Cov6 cc{cov_quadratic(wc, r0[0], r0[1], r0[2]),
                cov_bilinear(wc, r0[0], r0[1], r0[2], r1[0], r1[1], r1[2]),
                cov_bilinear(wc, r0[0], r0[1], r0[2], r2[0], r2[1], r2[2]),
                cov_quadratic(wc, r1[0], r1[1], r1[2]),
                cov_bilinear(wc, r1[0], r1[1], r1[2], r2[0], r2[1], r2[2]),
                cov_quadratic(wc, r2[0], r2[1], r2[2])};
  const float x_cl = fminf(fmaxf(x / z, -fp.lim_x), fp.lim_x) * z;
  const float y_cl = fminf(fmaxf(y / z, -fp.lim_y), fp.lim_y) * z;
  const float inv_z2 = inv_z * inv_z;
  const float j00 = fp.fx * inv_z;
  const float j02 = -fp.fx * x_cl * inv_z2;
  const float j11 = fp.fy * inv_z;
  const float j12 = -fp.fy * y_cl * inv_z2;
  const float b = j00 * j11 * cc.xy + j00 * j12 * cc.xz + j02 * j11 * cc.yz + j02 * j12 * cc.zz;
  float a = j00 * j00 * cc.xx + 2.0f * j00 * j02 * cc.xz + j02 * j02 * cc.zz;
  float c = j11 * j11 * cc.yy + 2.0f * j11 * j12 * cc.yz + j12 * j12 * cc.zz;
  a = a + fp.blur_x;
  c = c + fp.blur_y;
  const float apco2 = (a + c) * 0.5f;
  const float amco2 = (a - c) * 0.5f;
  const float term = sqrtf(fmaxf(amco2 * amco2 + b * b, 0.0f));
  const float major = fmaxf(apco2 + term, 1e-10f);
  const float minor = fmaxf(apco2 - term, 1e-10f);
  float udx = 1.0f, udy = 0.0f;
  if (fabsf(b) > 1e-8f) {
    // normalize2_or_default (wgsl:79-85), inverseSqrt as 1/sqrt.
    const float px = b, py = major - a;
    const float len2 = px * px + py * py;
    if (len2 > 1e-20f) {
      const float inv = 1.0f / sqrtf(len2);
      udx = px * inv;
      udy = py * inv;
    }
  } else if (a < c) {
    udx = 0.0f;
    udy = 1.0f;
  }
  float major_r = fminf(fmaxf(sqrtf(major) * 3.0f, 1e-4f), 2.0f);
  const float minor_r = fminf(fmaxf(sqrtf(minor) * 3.0f, 1e-4f), 2.0f);
  major_r = fminf(major_r, minor_r * 64.0f);
  const float ux = udx * major_r, uy = udy * major_r;
  const float vx = -udy * minor_r, vy = udx * minor_r;
  const float ext_x = fabsf(ux) + fabsf(vx);
  const float ext_y = fabsf(uy) + fabsf(vy);
  if (x_ndc + ext_x < -1.0f || x_ndc - ext_x > 1.0f || y_ndc + ext_y < -1.0f || y_ndc - ext_y > 1.0f) return o;
  o.drawn = true;
  o.cx = x_ndc;
  o.cy = y_ndc;
  o.ux = ux;
  o.uy = uy;
  o.vx = vx;
  o.vy = vy;
  o.alpha = src.alpha();
  // normalize3_or_default(position - camera) (wgsl:87-93, 329).
  float dxw = 0.0f, dyw = 0.0f, dzw = 1.0f;
  const float len2 = relx * relx + rely * rely + relz * relz;
  if (len2 > 1e-20f) {
    const float inv = 1.0f / sqrtf(len2);
    dxw = relx * inv;
    dyw = rely * inv;
    dzw = relz * inv;
  }
#pragma unroll
  for (uint32_t ch = 0; ch < 3; ++ch) {
    const float s = sh_channel<DEG>(src, ch, dxw, dyw, dzw);
    o.col[ch] = fminf(fmaxf(s + 0.5f, 0.0f), 1.0f) * o.alpha;
  }
  return o;
}

// Zeroes the word ranges of `c` with the whole grid (grid-stride).
__device__ __forceinline__ void clear_ranges(const Clears &c) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
  for (int r = 0; r < Clears::kRanges; ++r)
    for (size_t i = i0; i < c.words[r]; i += stride) c.ptr[r][i] = 0u;
}

// Padding slot g of view v: never visible, sorts last in its view.
__device__ __forceinline__ void write_padding(uint32_t v, uint32_t g, uint32_t *__restrict__ keys,
                                              uint32_t *__restrict__ vals, uint2 *__restrict__ rect) {
  keys[g] = 0u;
  vals[g] = (v << kLocalBits) | kLocalMask;
  rect[g] = make_uint2(0xFFFFFFFFu, 0u);
}

// Slot g of view v (frame fp) for its splat i, projected as p: depth key,
// value, 48-byte record and pixel bounds.
template <class Src>
__device__ __forceinline__ void write_slot(const FrameParams &fp, uint32_t v, uint32_t g, uint32_t i,
                                           const Projected &p, const Src &src, uint32_t *__restrict__ keys,
                                           uint32_t *__restrict__ vals, float4 *__restrict__ rec,
                                           uint2 *__restrict__ rect) {
  keys[g] = p.key;
  vals[g] = (v << kLocalBits) | i;
  uint2 rr = make_uint2(0xFFFFFFFFu, 0u);
  if (p.drawn) {
    const float cut = fminf(src.cut(), 1.0f);
    const float det = p.ux * p.vy - p.vx * p.uy;
    rec[3 * (size_t)g + 0] = make_float4(p.cx, p.cy, p.vy / det, -p.vx / det);
    rec[3 * (size_t)g + 1] = make_float4(-p.uy / det, p.ux / det, cut, p.alpha);
    rec[3 * (size_t)g + 2] = make_float4(p.col[0], p.col[1], p.col[2], 0.0f);
    if (cut >= 0.0f) {
      // Bounding box of {center + A l : |l|^2 <= cut} plus one pixel.
      const float sc = sqrtf(cut);
      const float ex = sc * sqrtf(p.ux * p.ux + p.vx * p.vx);
      const float ey = sc * sqrtf(p.uy * p.uy + p.vy * p.vy);
      const float hw = 0.5f * (float)fp.width, hh = 0.5f * (float)fp.height;
      const float cxp = (p.cx + 1.0f) * hw;
      const float cyp = (1.0f - p.cy) * hh;
      const float exw = ex * hw + 1.0f;
      const float eyh = ey * hh + 1.0f;
      const float x0f = ceilf(cxp - exw - 0.5f), x1f = floorf(cxp + exw - 0.5f);
      const float y0f = ceilf(cyp - eyh - 0.5f), y1f = floorf(cyp + eyh - 0.5f);
      const float wm1 = (float)(fp.width - 1), hm1 = (float)(fp.height - 1);
      if (!(x1f < 0.0f || y1f < 0.0f || x0f > wm1 || y0f > hm1 || x0f > x1f || y0f > y1f)) {
        const uint32_t x0 = (uint32_t)fmaxf(x0f, 0.0f), x1 = (uint32_t)fminf(x1f, wm1);
        const uint32_t y0 = (uint32_t)fmaxf(y0f, 0.0f), y1 = (uint32_t)fminf(y1f, hm1);
        rr = make_uint2(x0 | (y0 << 16), x1 | (y1 << 16));
      }
    }
  }
  rect[g] = rr;
}

// The workgroup's smallest non-zero key (~0u when none) and largest key,
// stored at range[blockIdx.x].
__device__ __forceinline__ void store_key_range(uint32_t key, uint2 *__restrict__ range) {
  __shared__ uint32_t part[2][kWaves];
  uint32_t mn = key ? key : 0xFFFFFFFFu, mx = key;
#pragma unroll
  for (uint32_t d = 1; d < kWave; d <<= 1) {
    mn = min(mn, (uint32_t)__shfl_xor((int)mn, (int)d, kWave));
    mx = max(mx, (uint32_t)__shfl_xor((int)mx, (int)d, kWave));
  }
  const uint32_t w = threadIdx.x / kWave;
  if (__lane_id() == 0) {
    part[0][w] = mn;
    part[1][w] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (uint32_t q = 1; q < kWaves; ++q) {
      mn = min(mn, part[0][q]);
      mx = max(mx, part[1][q]);
    }
    range[blockIdx.x] = make_uint2(min(mn, part[0][0]), max(mx, part[1][0]));
  }
}

template <uint32_t FMT, bool INLINE>
__global__ __launch_bounds__(kThreads) void k_project(const ViewDev *__restrict__ views,
                                                      const uint32_t *__restrict__ seg_start, uint32_t nviews,
                                                      uint32_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                                      float4 *__restrict__ rec, uint2 *__restrict__ rect,
                                                      Clears clears, uint2 *__restrict__ key_range,
                                                      InlineTable table) {
  clear_ranges(clears);
  if (INLINE && blockIdx.x == 0)
    for (uint32_t i = threadIdx.x; i < table.count; i += kThreads) table.dst[i] = table.words[i];
  const uint32_t g = blockIdx.x * kThreads + threadIdx.x;
  const uint32_t v = INLINE ? 0u : find_view(seg_start, nviews, blockIdx.x * kThreads);
  const ViewDev &vw = INLINE ? *reinterpret_cast<const ViewDev *>(&table.words[kInlineViewWord]) : views[v];
  if (vw.format != FMT) return;  // uniform: one view per workgroup
  const uint32_t i = g - vw.seg_off;
  uint32_t key = 0;
  if (i >= vw.n) {
    write_padding(v, g, keys, vals, rect);
  } else {
    const auto src = FormatOf<FMT>::at(vw.soa, vw.n, i);
    const Projected p = project_one<FormatOf<FMT>::kDeg>(vw.fp, src);
    write_slot(vw.fp, v, g, i, p, src, keys, vals, rec, rect);
    key = p.key;
  }
  if (key_range) store_key_range(key, key_range);
}

// Views of one scene projected together: workgroup b takes chunk
// chunk_lo + b, i.e. splats [256 c, 256 c + 256) of the group whose chunk
// range holds it; each thread loads its splat's fields once and writes its
// slot in every view of the group (the slot layout and values are those of
// k_project, so the results are the same bit for bit).
template <uint32_t FMT>
__global__ __launch_bounds__(kThreads) void k_project_groups(const ViewDev *__restrict__ views,
                                                             const ProjGroup *__restrict__ groups,
                                                             uint32_t ngroups, uint32_t chunk_lo,
                                                             uint32_t *__restrict__ keys,
                                                             uint32_t *__restrict__ vals, float4 *__restrict__ rec,
                                                             uint2 *__restrict__ rect, Clears clears) {
  clear_ranges(clears);
  const uint32_t chunk = chunk_lo + blockIdx.x;
  uint32_t lo = 0, hi = ngroups;  // last group whose first chunk <= chunk
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (groups[mid].chunk_start <= chunk) lo = mid; else hi = mid;
  }
  const ProjGroup &gr = groups[lo];
  const uint32_t i = (chunk - gr.chunk_start) * kThreads + threadIdx.x;
  const uint32_t nv = gr.nv;
  if (i >= gr.n) {
    for (uint32_t q = 0; q < nv; ++q) {
      const uint32_t v = gr.views[q];
      write_padding(v, views[v].seg_off + i, keys, vals, rect);
    }
    return;
  }
  constexpr int DEG = FormatOf<FMT>::kDeg;
  RegSource<DEG> src;
  src.load(FormatOf<FMT>::at(gr.soa, gr.n, i));
  for (uint32_t q = 0; q < nv; ++q) {
    const uint32_t v = gr.views[q];
    const ViewDev &vw = views[v];
    write_slot(vw.fp, v, vw.seg_off + i, i, project_one<DEG>(vw.fp, src), src, keys, vals, rec, rect);
  }
}

template <uint32_t FMT>
__global__ __launch_bounds__(kThreads) void k_project_export(const ViewDev *__restrict__ views, uint32_t v,
                                                             uint32_t *__restrict__ keys_out,
                                                             float *__restrict__ export12) {
  const ViewDev &vw = views[v];
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= vw.n) return;
  const Projected p = project_one<FormatOf<FMT>::kDeg>(vw.fp, FormatOf<FMT>::at(vw.soa, vw.n, i));
  if (keys_out) keys_out[i] = p.key;
  float *o = export12 + 12 * (size_t)i;
  // invalid_splat() (wgsl:122-129): zeros for a splat that is not drawn.
  o[0] = p.drawn ? 1.0f : 0.0f;
  o[1] = p.cx;
  o[2] = p.cy;
  o[3] = p.drawn ? p.ux : 0.0f;
  o[4] = p.drawn ? p.uy : 0.0f;
  o[5] = p.drawn ? p.vx : 0.0f;
  o[6] = p.drawn ? p.vy : 0.0f;
  o[7] = p.col[0];
  o[8] = p.col[1];
  o[9] = p.col[2];
  o[10] = p.alpha;
  o[11] = 0.0f;
}

// ------------------------------------------------------------- tile binning

// Number of tiles of (16 << shift) pixels a pixel-bounds rectangle
// (x0|y0<<16, x1|y1<<16) touches.
__device__ __forceinline__ uint32_t rect_tiles(uint2 r, uint32_t shift) {
  if (r.x == 0xFFFFFFFFu) return 0u;
  const uint32_t sh = 4 + shift;
  const uint32_t x0 = (r.x & 0xFFFFu) >> sh, y0 = (r.x >> 16) >> sh;
  const uint32_t x1 = (r.y & 0xFFFFu) >> sh, y1 = (r.y >> 16) >> sh;
  return (x1 - x0 + 1) * (y1 - y0 + 1);
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long x) {
#pragma unroll
  for (uint32_t off = kWave / 2; off > 0; off >>= 1) {
    const uint32_t lo = __shfl_xor((uint32_t)x, off, kWave), hi = __shfl_xor((uint32_t)(x >> 32), off, kWave);
    x += ((unsigned long long)hi << 32) | lo;
  }
  return x;
}

// Tile counts in reversed slot order with their in-block exclusive scan.
// Block b covers reversed positions [1024 b, 1024 b + 1024), i.e. sorted
// slots total-1-r, four consecutive per thread (a block lies in one view
// segment, since segments align to 4096 slots).
__global__ __launch_bounds__(kThreads) void k_tile_counts(const ViewDev *__restrict__ views,
                                                          const uint32_t *__restrict__ seg_start, uint32_t nviews,
                                                          uint32_t total,
                                                          const uint32_t *__restrict__ skeys,
                                                          const uint32_t *__restrict__ svals,
                                                          const uint2 *__restrict__ rect,
                                                          uint32_t *__restrict__ counts_rev,
                                                          uint32_t *__restrict__ offsets_rev,
                                                          uint32_t *__restrict__ visible,
                                                          uint32_t *__restrict__ block_sums,
                                                          uint32_t *__restrict__ block_pairs,
                                                          uint32_t *__restrict__ rect_lo,
                                                          uint32_t *__restrict__ rect_hi, Clears clears) {
  __shared__ uint32_t lds4[kWaves];
  __shared__ uint32_t wfine[kWaves];
  clear_ranges(clears);
  const uint32_t t = threadIdx.x, w = t / kWave, lane = __lane_id();
  // Workgroups reach the XCDs round robin (grid a multiple of kXcds): each
  // XCD takes one contiguous run of blocks in order, so the rect gathers of
  // a view's blocks meet in one L2 instead of in all eight.
  const uint32_t b = (blockIdx.x % kXcds) * (gridDim.x / kXcds) + blockIdx.x / kXcds;
  if (b >= total / kCountBlock) return;
  const uint32_t slot_lo = total - kCountBlock * (b + 1);  // lowest slot of the block
  const uint32_t v = find_view(seg_start, nviews, slot_lo);
  const ViewDev &vw = views[v];
  const uint32_t shift = vw.bin_shift;
  const uint32_t r0 = b * kCountBlock + t * 4;
  // Loads first (keys, values, then the rects they name), so that each
  // level's memory latency is paid once rather than once per item.
  uint32_t key[4], local[4];
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    const uint32_t g = total - 1 - (r0 + k);
    key[k] = skeys[g];
    local[k] = svals[g] - (v << kLocalBits);
  }
  // The key after item k in slot order is item k-1's; after item 0 it is
  // the previous thread's (or block's) last item.
  const uint32_t g_first = total - 1 - r0;
  const uint32_t next0 = g_first + 1 < total ? skeys[g_first + 1] : 0u;
  const uint32_t n_v = vw.n, off_v = vw.seg_off;
  bool live[4];
  uint2 rc[4];
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    live[k] = key[k] != 0u && local[k] < n_v;
    rc[k] = rect[live[k] ? off_v + local[k] : 0u];  // unconditional: the four loads issue together
  }
  // The rects in sorted slot order, for the pair emission to read in order.
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    const uint32_t g = total - 1 - (r0 + k);
    rect_lo[g] = rc[k].x;
    rect_hi[g] = rc[k].y;
  }
  uint32_t c[4];
  uint32_t sum = 0, fine = 0;
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    uint32_t cc = 0;
    if (live[k]) {
      cc = rect_tiles(rc[k], shift);
      fine += rect_tiles(rc[k], 0);
      const uint32_t j = total - 1 - (r0 + k) - off_v;
      const uint32_t next = k == 0 ? next0 : key[k - 1];
      if (j + 1 == vw.seg_len || next == 0u) visible[v] = j + 1;
    }
    c[k] = cc;
    sum += cc;
  }
  uint32_t block_total;
  uint32_t run = block_exclusive_scan(sum, lds4, &block_total);
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    counts_rev[r0 + k] = c[k];
    offsets_rev[r0 + k] = run;
    run += c[k];
  }
#pragma unroll
  for (uint32_t off = kWave / 2; off > 0; off >>= 1) fine += __shfl_down(fine, off, kWave);
  if (lane == 0) wfine[w] = fine;
  __syncthreads();
  if (t == 0) {
    block_sums[b] = block_total;
    block_pairs[b] = wfine[0] + wfine[1] + wfine[2] + wfine[3];
  }
}

// One workgroup: exclusive scan of the block totals in place (each block's
// first pair index), the binning pair total M (info[1..2]), its overflow
// flag (info[3]) and *m_out = min(M, capacity), and the sum of block_pairs
// into *raster_pairs. Thread t of the 1024 owns the contiguous block range
// [t * per, (t + 1) * per): it sums its blocks, the workgroup scans the
// thread sums (64-bit), and each thread writes its blocks' exclusive prefix.
constexpr uint32_t kBlockScanThreads = 1024;

__global__ __launch_bounds__(kBlockScanThreads) void k_block_scan(uint32_t *__restrict__ block_sums, uint32_t blocks,
                                                                  uint32_t capacity, uint32_t *__restrict__ m_out,
                                                                  uint32_t *__restrict__ info,
                                                                  const uint32_t *__restrict__ block_pairs,
                                                                  unsigned long long *__restrict__ raster_pairs) {
  constexpr uint32_t W = kBlockScanThreads / kWave;
  __shared__ unsigned long long wsum[W], wpairs[W];
  const uint32_t t = threadIdx.x, lane = __lane_id(), w = t / kWave;
  const uint32_t per = (blocks + kBlockScanThreads - 1) / kBlockScanThreads;
  const uint32_t b0 = min(t * per, blocks), b1 = min(b0 + per, blocks);
  unsigned long long sum = 0, pairs = 0;
#pragma unroll 8
  for (uint32_t i = b0; i < b1; ++i) {
    sum += block_sums[i];
    pairs += block_pairs[i];
  }
  unsigned long long inc = sum;
#pragma unroll
  for (uint32_t d = 1; d < kWave; d <<= 1) {
    const unsigned long long y = __shfl_up(inc, d, kWave);
    if (lane >= d) inc += y;
  }
  pairs = wave_sum_u64(pairs);
  if (lane == kWave - 1) wsum[w] = inc;
  if (lane == 0) wpairs[w] = pairs;
  __syncthreads();
  unsigned long long run = inc - sum, total = 0, total_pairs = 0;
#pragma unroll
  for (uint32_t q = 0; q < W; ++q) {
    run += q < w ? wsum[q] : 0ull;
    total += wsum[q];
    total_pairs += wpairs[q];
  }
#pragma unroll 8
  for (uint32_t i = b0; i < b1; ++i) {
    const uint32_t x = block_sums[i];
    block_sums[i] = (uint32_t)run;
    run += x;
  }
  if (t != 0) return;
  *raster_pairs = total_pairs;
  const uint64_t m = total;
  *m_out = m > capacity ? capacity : (uint32_t)m;
  info[1] = (uint32_t)m;
  info[2] = (uint32_t)(m >> 32);
  info[3] = m > capacity ? 1u : 0u;
}

// Load-balanced expansion. Workgroup b owns sorted slots [256b, 256b+256) of
// one view; their pairs are one contiguous range of the output, nearest slot
// first. The slots' first pair index, tile rectangle and record slot are
// staged in LDS in pair order; then each pass writes 256 consecutive pairs,
// every thread finding its pair's slot by binary search.
__global__ __launch_bounds__(kThreads) void k_tile_emit(const ViewDev *__restrict__ views,
                                                        const uint32_t *__restrict__ seg_start, uint32_t nviews,
                                                        uint32_t total,
                                                        const uint32_t *__restrict__ svals,
                                                        const uint32_t *__restrict__ rect_lo,
                                                        const uint32_t *__restrict__ rect_hi,
                                                        const uint32_t *__restrict__ counts_rev,
                                                        const uint32_t *__restrict__ offsets_rev,
                                                        const uint32_t *__restrict__ block_sums, uint32_t capacity,
                                                        uint32_t *__restrict__ pair_keys,
                                                        uint32_t *__restrict__ pair_vals, PairHist ph,
                                                        EmitTotals et) {
  __shared__ uint32_t s_hist[kPairHistPasses][1u << 9];
  __shared__ uint32_t s_first[kThreads + 1];
  __shared__ uint32_t s_slot[kThreads];
  __shared__ uint32_t s_x0[kThreads], s_y0[kThreads], s_w[kThreads];
  __shared__ uint32_t s_red[kWaves];
  __shared__ unsigned long long s_red64[2][kWaves];
  const uint32_t g0 = blockIdx.x * kThreads;
  const uint32_t v = find_view(seg_start, nviews, g0);
  const ViewDev &vw = views[v];
  const uint32_t t = threadIdx.x, w = t / kWave, lane = __lane_id();
  // Staged entry t is sorted slot g0 + 255 - t (nearest first); the
  // workgroup's 256 entries lie in one count block.
  const uint32_t r = total - 1 - (g0 + kThreads - 1 - t);
  const uint32_t rb = (total - g0 - kThreads) / kCountBlock;
  uint32_t block_first;
  if (et.blocks) {  // raw block totals: sum those before this block
    uint32_t part = 0;
    for (uint32_t i = t; i < rb; i += kThreads) part += block_sums[i];
    (void)block_exclusive_scan(part, s_red, &block_first);
    if (blockIdx.x == 0) {  // what the block scan writes
      unsigned long long m = 0, pairs = 0;
      for (uint32_t i = t; i < et.blocks; i += kThreads) {
        m += block_sums[i];
        pairs += et.block_pairs[i];
      }
      m = wave_sum_u64(m);
      pairs = wave_sum_u64(pairs);
      if (lane == 0) {
        s_red64[0][w] = m;
        s_red64[1][w] = pairs;
      }
      __syncthreads();
      if (t == 0) {
        m = pairs = 0;
#pragma unroll
        for (uint32_t q = 0; q < kWaves; ++q) {
          m += s_red64[0][q];
          pairs += s_red64[1][q];
        }
        *et.raster_pairs = pairs;
        *et.m_out = m > capacity ? capacity : (uint32_t)m;
        et.info[1] = (uint32_t)m;
        et.info[2] = (uint32_t)(m >> 32);
        et.info[3] = m > capacity ? 1u : 0u;
      }
    }
  } else {
    block_first = block_sums[rb];
  }
  const uint32_t cnt = counts_rev[r];
  const uint32_t first = block_first + offsets_rev[r];
  const uint32_t g = g0 + kThreads - 1 - t;
  const uint32_t slot = vw.seg_off + (svals[g] - (v << kLocalBits));  // loaded beside cnt
  const uint2 b = make_uint2(rect_lo[g], rect_hi[g]);                 // in slot order
  const uint32_t dmask = (1u << ph.digit_bits) - 1u;
  if (ph.hist)
    for (uint32_t i = t; i < ph.passes << ph.digit_bits; i += kThreads) s_hist[i >> ph.digit_bits][i & dmask] = 0;
  s_first[t] = first;
  if (t == kThreads - 1) s_first[kThreads] = first + cnt;
  if (cnt) {
    const uint32_t sh = 4 + vw.bin_shift;
    const uint32_t x0 = (b.x & 0xFFFFu) >> sh, x1 = (b.y & 0xFFFFu) >> sh;
    s_slot[t] = slot;
    s_x0[t] = x0;
    s_y0[t] = (b.x >> 16) >> sh;
    s_w[t] = x1 - x0 + 1;
  }
  __syncthreads();
  const uint32_t p_begin = s_first[0], p_end = min(s_first[kThreads], capacity);
  const uint32_t tiles_x = vw.fp.ctiles_x;
  for (uint32_t p = p_begin + t; p < p_end; p += kThreads) {
    // Largest i with s_first[i] <= p: the entry whose range holds p (entries
    // without pairs share the next entry's first index and are passed over).
    uint32_t lo = 0, hi = kThreads - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) / 2;
      if (s_first[mid] <= p)
        lo = mid;
      else
        hi = mid - 1;
    }
    const uint32_t k = p - s_first[lo];
    const uint32_t w = s_w[lo];
    const uint32_t tile = (s_y0[lo] + k / w) * tiles_x + s_x0[lo] + k % w;
    pair_keys[p] = (v << kPairViewShift) | tile;
    pair_vals[p] = s_slot[lo];
    if (ph.hist)
      for (uint32_t q = 0; q < ph.passes; ++q) atomicAdd(&s_hist[q][(tile >> ph.shift[q]) & dmask], 1u);
  }
  if (ph.hist) {
    __syncthreads();
    uint32_t *dst = ph.hist + (size_t)(blockIdx.x % kSortStripes) * kSortMaxPasses * (dmask + 1);
    for (uint32_t i = t; i < ph.passes << ph.digit_bits; i += kThreads) {
      const uint32_t c = s_hist[i >> ph.digit_bits][i & dmask];
      if (c) atomicAdd(&dst[i], c);
    }
  }
}

// The pairs are sorted by local tile only; each view's pairs were emitted
// as one contiguous run, so every (view, local tile) key is one contiguous
// run of the sorted pairs.
// The grid first clears `clears`.
__global__ __launch_bounds__(kThreads) void k_tile_ranges(const ViewDev *__restrict__ views,
                                                          const uint32_t *__restrict__ keys,
                                                          const uint32_t *__restrict__ m_dev,
                                                          uint2 *__restrict__ ranges, Clears clears,
                                                          WordCopy copy) {
  clear_ranges(clears);
  if (blockIdx.x == 0)
    for (uint32_t i = threadIdx.x; i < copy.words; i += kThreads) copy.dst[i] = copy.src[i];
  const uint32_t m = *m_dev;
  // Four keys per thread and step (one 16-byte load); the view table is read
  // only at a run boundary, once per binning tile.
  const uint32_t quads = m / 4;
  for (uint32_t q = blockIdx.x * kThreads + threadIdx.x; q < quads; q += gridDim.x * kThreads) {
    const uint4 k4 = reinterpret_cast<const uint4 *>(keys)[q];
    const uint32_t i = 4 * q;
    const uint32_t prev = i == 0 ? ~k4.x : keys[i - 1];
    const uint32_t next = i + 4 == m ? ~k4.w : keys[i + 4];
    const uint32_t k[4] = {k4.x, k4.y, k4.z, k4.w};
    const uint32_t before[4] = {prev, k4.x, k4.y, k4.z};
    const uint32_t after[4] = {k4.y, k4.z, k4.w, next};
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
      const bool first = before[u] != k[u], last = after[u] != k[u];
      if (first || last) {
        const uint32_t t = views[k[u] >> kPairViewShift].ctile_off + (k[u] & kPairTileMask);
        if (first) ranges[t].x = i + u;
        if (last) ranges[t].y = i + u + 1;
      }
    }
  }
  for (uint32_t i = 4 * quads + blockIdx.x * kThreads + threadIdx.x; i < m; i += gridDim.x * kThreads) {
    const uint32_t k = keys[i];
    const bool first = i == 0 || keys[i - 1] != k, last = i == m - 1 || keys[i + 1] != k;
    if (first || last) {
      const uint32_t t = views[k >> kPairViewShift].ctile_off + (k & kPairTileMask);
      if (first) ranges[t].x = i;
      if (last) ranges[t].y = i + 1;
    }
  }
}

// Heaviest binning tiles first: binning tile c (global) goes to class
// clamp(12 - floor(log2 list length), 0, 7) (class 0: 4096 or more pairs,
// class 7: fewer than 64 or none) and takes the next place of its class,
// order[class * total_ctiles + place]; counts[class] ends as the class size
// (zero before). The raster walks the classes in turn.
__global__ __launch_bounds__(kThreads) void k_raster_order(const uint2 *__restrict__ ranges, uint32_t total_ctiles,
                                                           uint32_t *__restrict__ counts,
                                                           uint32_t *__restrict__ order) {
  __shared__ uint32_t local[kOrderClasses], base[kOrderClasses];
  const uint32_t t = threadIdx.x;
  if (t < kOrderClasses) local[t] = 0;
  __syncthreads();
  const uint32_t c = blockIdx.x * kThreads + t;
  uint32_t cls = 0, place = 0;
  if (c < total_ctiles) {
    const uint2 r = ranges[c];
    cls = raster_order_class(r.y - r.x);
    place = atomicAdd(&local[cls], 1u);
  }
  __syncthreads();
  if (t < kOrderClasses) base[t] = local[t] ? atomicAdd(&counts[t], local[t]) : 0u;
  __syncthreads();
  if (c < total_ctiles) order[(size_t)cls * total_ctiles + base[cls] + place] = c;
}

// One thread per slot of view v: every visible splat adds one to each 16x16
// tile its pixel bounds overlap.
__global__ __launch_bounds__(kThreads) void k_raster_tile_counts(const ViewDev *__restrict__ views, uint32_t v,
                                                                 const uint32_t *__restrict__ skeys,
                                                                 const uint32_t *__restrict__ svals,
                                                                 const uint2 *__restrict__ rect,
                                                                 uint32_t *__restrict__ counts) {
  const ViewDev &vw = views[v];
  const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= vw.seg_len) return;
  const uint32_t g = vw.seg_off + j;
  const uint32_t local = svals[g] - (v << kLocalBits);
  if (skeys[g] == 0u || local >= vw.n) return;
  const uint2 r = rect[vw.seg_off + local];
  if (r.x == 0xFFFFFFFFu) return;
  const uint32_t x0 = (r.x & 0xFFFFu) / kTile, y0 = (r.x >> 16) / kTile;
  const uint32_t x1 = (r.y & 0xFFFFu) / kTile, y1 = (r.y >> 16) / kTile;
  for (uint32_t ty = y0; ty <= y1; ++ty)
    for (uint32_t tx = x0; tx <= x1; ++tx) atomicAdd(&counts[ty * vw.fp.tiles_x + tx], 1u);
}

__global__ void k_fill_u32(uint32_t *__restrict__ p, uint32_t value, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

// ------------------------------------------------------------------ raster

// Work a raster launch may take over from the binning kernels it follows:
// zeroing scratch words for the next frame (clears) and handing the frame
// summary to pinned host memory (copy, by workgroup 0). Neither touches
// what the raster reads.
__device__ __forceinline__ void raster_prologue(const Clears &clears, const WordCopy &copy) {
  // 64 words per workgroup, taken by the workgroups dispatched last (the
  // lightest tiles), so that the heaviest tiles start without the stores
  const uint32_t rb = gridDim.x - 1 - blockIdx.x;
  if (threadIdx.x < 64)
    for (int r = 0; r < Clears::kRanges; ++r)
      for (size_t i = (size_t)rb * 64 + threadIdx.x; i < clears.words[r]; i += (size_t)gridDim.x * 64)
        clears.ptr[r][i] = 0u;
  if (blockIdx.x == 0)
    for (uint32_t i = threadIdx.x; i < copy.words; i += kThreads) copy.dst[i] = copy.src[i];
}

// Records a raster stage holds: a blend starts once blend_min are staged, so
// at most blend_min - 1 are left when up to kThreads more are added; the
// blend pads the stage to a multiple of four.
__host__ __device__ __forceinline__ uint32_t stage_capacity(uint32_t blend_min) {
  return (blend_min + kThreads - 1 + 3) & ~3u;
}

// Raster workgroup order, in one of two layouts.
// Per view (order null): view v owns the workgroups [rblock_start[v],
// rblock_start[v] + raster_blocks) (its binning tiles in whole runs of kXcds,
// times S = 4^shift raster tiles each, shift the view's bin shift). Its local
// workgroup l takes binning tile c = (l / (kXcds S)) kXcds + l % kXcds and
// raster tile k = (l / kXcds) % S of it.
// Ordered (every view at one bin shift): workgroup b takes place
// j = (b / (kXcds S)) kXcds + b % kXcds of the heaviest-first order of all
// binning tiles (k_raster_order) and raster tile k = (b / kXcds) % S, so the
// long lists start first and the kernel does not end on them.
// Either way the S raster tiles of one binning tile run back to back on one
// XCD (workgroups are dealt to the XCDs round robin) and find their shared
// list's records in its L2. False for padding (no such binning tile, or a
// raster tile past the view's edge).
__device__ __forceinline__ bool raster_tile(const ViewDev *__restrict__ views,
                                            const uint32_t *__restrict__ rblock_start, uint32_t nviews,
                                            uint32_t total_rblocks, const RasterOrder &ro, uint32_t b, uint32_t &v,
                                            uint32_t &tx, uint32_t &ty) {
  if (b >= total_rblocks) return false;
  uint32_t c, k, shift;
  if (ro.order) {
    shift = ro.shift;
    const uint32_t S = 1u << (2 * shift);
    const uint32_t j = (b / (kXcds * S)) * kXcds + b % kXcds;
    k = (b / kXcds) % S;
    if (j >= ro.total_ctiles) return false;
    uint32_t cls = 0, first = 0;
    for (; cls + 1 < kOrderClasses; ++cls) {
      const uint32_t n = ro.counts[cls];
      if (j < first + n) break;
      first += n;
    }
    const uint32_t g = ro.order[(size_t)cls * ro.total_ctiles + (j - first)];
    uint32_t lo = 0, hi = nviews - 1;  // last view whose first binning tile <= g
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) / 2;
      if (views[mid].ctile_off <= g) lo = mid; else hi = mid - 1;
    }
    v = __builtin_amdgcn_readfirstlane(lo);
    c = g - views[v].ctile_off;
  } else {
    v = find_view(rblock_start, nviews, b);
    shift = views[v].bin_shift;
    const uint32_t S = 1u << (2 * shift);
    const uint32_t l = b - views[v].rblock_off;
    c = (l / (kXcds * S)) * kXcds + l % kXcds;
    k = (l / kXcds) % S;
  }
  const ViewDev &vw = views[v];
  if (c >= vw.ctiles) return false;
  tx = ((c % vw.fp.ctiles_x) << shift) + (k & ((1u << shift) - 1u));
  ty = ((c / vw.fp.ctiles_x) << shift) + (k >> shift);
  return tx < vw.fp.tiles_x && ty < vw.fp.tiles_y;
}

// Final pixel (x, y) of view vw: RGBA8 (rounded once, clamped) and RGBA32F
// in the device images, or RGBA8 alone into the view's host target.
__device__ __forceinline__ void write_pixel(const ViewDev &vw, uint32_t x, uint32_t y, float cr, float cg, float cb,
                                            float T, float4 *__restrict__ img32, uchar4 *__restrict__ img8) {
  const float A = 1.0f - T;
  auto q = [](float val) { return (unsigned char)(fminf(fmaxf(val, 0.0f), 1.0f) * 255.0f + 0.5f); };
  const uchar4 p8 = make_uchar4(q(cr), q(cg), q(cb), q(A));
  const size_t pix = (size_t)y * vw.fp.width + x;
  if (vw.out8) {
    vw.out8[pix] = p8;
    return;
  }
  img32[vw.img_off + pix] = make_float4(cr, cg, cb, A);
  img8[vw.img_off + pix] = p8;
}

// Blends staged splats [0, n) of the LDS stage into the thread's pixel.
// LOOP 0: per-lane exit at the early stop, divergent blend branch.
// LOOP 1: wave-uniform loop, four splats per wave liveness check; the stage
// is padded to a multiple of four with splats no pixel can hit (cut = -1).
// Both make the same operations in the same order for every pixel.
// True unless the splat's coverage ellipse {p : |L (p - c)|^2 <= cut}
// (record halves a = [cx, cy, L00, L01], b = [L10, L11, cut, alpha], NDC)
// misses the rectangle [xlo, xhi] x [ylo, yhi]. With c outside the
// rectangle the convex quadratic's minimum lies on an edge, where it is the
// clamped 1-D minimum. Used with the rectangle of a tile's pixel edges (half
// a pixel beyond its pixel centres) and a 1% slack on cut, so a splat it
// rejects passes the coverage test at none of the tile's pixels.
__device__ __forceinline__ bool ellipse_meets_rect(const float4 &a, const float4 &b, float xlo, float xhi, float ylo,
                                                   float yhi) {
  const float cx = a.x, cy = a.y;
  if (cx >= xlo && cx <= xhi && cy >= ylo && cy <= yhi) return true;
  const float l00 = a.z, l01 = a.w, l10 = b.x, l11 = b.y;
  const float A = l00 * l00 + l10 * l10, B = l00 * l01 + l10 * l11, C = l01 * l01 + l11 * l11;
  const float ra = A > 0.0f ? __builtin_amdgcn_rcpf(A) : 0.0f;
  const float rc = C > 0.0f ? __builtin_amdgcn_rcpf(C) : 0.0f;
  const float dxlo = xlo - cx, dxhi = xhi - cx, dylo = ylo - cy, dyhi = yhi - cy;
  auto q = [&](float dx, float dy) {
    const float lx = l00 * dx + l01 * dy, ly = l10 * dx + l11 * dy;
    return lx * lx + ly * ly;
  };
  const float m0 = q(dxlo, fminf(fmaxf(-B * dxlo * rc, dylo), dyhi));
  const float m1 = q(dxhi, fminf(fmaxf(-B * dxhi * rc, dylo), dyhi));
  const float m2 = q(fminf(fmaxf(-B * dylo * ra, dxlo), dxhi), dylo);
  const float m3 = q(fminf(fmaxf(-B * dyhi * ra, dxlo), dxhi), dyhi);
  return fminf(fminf(m0, m1), fminf(m2, m3)) <= b.z * 1.01f + 1e-6f;
}

// One passing evaluation of a staged splat at a pixel, the same operations
// in every raster loop: g = exp(-4.5 r2) evaluated as 2^(r2 * -4.5 log2 e)
// (one multiply and the hardware exp2), weight w = T g, colour += w times
// the premultiplied rgb, T -= w alpha (T (1 - alpha g) up to rounding). The
// reference's back-to-front "over" equals this front-to-back form up to the
// T < 1e-4 early stop (draw_pass.rs:75).
__device__ __forceinline__ void blend_hit(float r2, float alpha, const float4 &c, float &T, float &cr, float &cg,
                                          float &cb) {
  constexpr float kNeg45Log2e = -6.4921276840f;  // -4.5 / ln 2
  const float g = __builtin_amdgcn_exp2f(r2 * kNeg45Log2e);
  const float w = T * g;
  cr = fmaf(w, c.x, cr);
  cg = fmaf(w, c.y, cg);
  cb = fmaf(w, c.z, cb);
  T = fmaf(-w, alpha, T);
}

template <int LOOP, bool COUNT>
__device__ __forceinline__ void blend_stage(const float4 *s0, float4 *s1, const float4 *s2, uint32_t n, float pnx,
                                            float pny, float &T, float &cr, float &cg, float &cb, bool &done,
                                            uint32_t &passed, uint32_t &tested) {
  if (LOOP == 0) {
    if (done) return;
    for (uint32_t j = 0; j < n; ++j) {
      const float4 a = s0[j];
      const float4 b = s1[j];
      const float dx = pnx - a.x;
      const float dy = pny - a.y;
      const float lx = fmaf(a.w, dy, a.z * dx);
      const float ly = fmaf(b.y, dy, b.x * dx);
      const float r2 = fmaf(ly, ly, lx * lx);
      if (COUNT) ++tested;
      if (r2 <= b.z) {
        if (COUNT) ++passed;
        blend_hit(r2, b.w, s2[j], T, cr, cg, cb);
        if (T < 1e-4f) {
          done = true;
          return;
        }
      }
    }
  } else {
    constexpr uint32_t U = 4;
    const uint32_t nu = (n + U - 1) & ~(U - 1);
    const uint32_t t = threadIdx.x;
    if (t < nu - n) s1[n + t] = make_float4(0.0f, 0.0f, -1.0f, 0.0f);
    __syncthreads();
    bool active = !done;
    for (uint32_t j0 = 0; j0 < nu; j0 += U) {
      if (__ballot(active) == 0) break;  // wave-uniform
      // the group's coverage terms first: one LDS latency per group
      float4 ga[U], gb[U], gc[U];
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        ga[u] = s0[j0 + u];
        gb[u] = s1[j0 + u];
        gc[u] = s2[j0 + u];
      }
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        const uint32_t j = j0 + u;
        const float4 a = ga[u];
        const float4 b = gb[u];
        const float dx = pnx - a.x;
        const float dy = pny - a.y;
        const float lx = fmaf(a.w, dy, a.z * dx);
        const float ly = fmaf(b.y, dy, b.x * dx);
        const float r2 = fmaf(ly, ly, lx * lx);
        const bool hit = active && r2 <= b.z;
        if (COUNT) {
          tested += (active && j < n) ? 1u : 0u;
          passed += hit ? 1u : 0u;
        }
        if (hit) {
          blend_hit(r2, b.w, gc[u], T, cr, cg, cb);
          active = !(T < 1e-4f);
        }
      }
    }
    done = !active;
  }
}

// One workgroup per 16x16 tile, one thread per pixel. The workgroup walks
// its binning tile's list (depth order) 256 entries at a time; the entries
// whose pixel bounds overlap the tile are compacted in order into the LDS
// stage (records gathered from global memory), and the stage is blended
// front to back once it holds blend_min splats or the list ends (blend_hit:
// C += T*g*rgb_a, T -= T*g*a), stopping when T < 1e-4 (equivalent to the
// reference's back-to-front premultiplied "over", draw_pass.rs:75, up to that
// bound).
// Coverage is r^2 <= cut (the fragment test of wgsl:358-367 precomputed per
// splat); g = exp(-4.5 r^2) (wgsl:363).
// COUNT: also adds to evals[0] the number of (pixel, splat) evaluations that
// pass coverage before the pixel's early stop (the raster's algorithmic work)
// and to evals[1] the number of coverage tests made.
// The work of raster workgroup b (a 16x16 tile, see raster_tile) by the
// calling workgroup, with the LDS stage s0/s1/s2 and the per-wave counts
// wkeep.
// 8x8 raster tiles (raster variant Blocks8): one 8x8 pixel block per wave.
// The workgroup stages the records of its 16x16 tile's kept splats once
// (s0/s1/s2) and keeps, for each block, the list of staged positions whose
// pixel bounds meet the block (lists after s2, cap words each); wave w blends
// only block w's list with the wave-uniform loop. A splat left out of a
// block's list covers none of its pixels (the bounds are conservative), so
// every pixel blends the same splats in the same order as in the other
// variants.
template <bool COUNT>
__device__ __forceinline__ void raster_workgroup_blocks(uint32_t b, const ViewDev *__restrict__ views,
                                                        const uint32_t *__restrict__ rblock_start, uint32_t nviews,
                                                        uint32_t total_rblocks, const RasterOrder &ro,
                                                        uint32_t blend_min, const float4 *__restrict__ rec,
                                                        const uint2 *__restrict__ rect,
                                                        const uint2 *__restrict__ ranges,
                                                        const uint32_t *__restrict__ list, float4 *__restrict__ img32,
                                                        uchar4 *__restrict__ img8,
                                                        unsigned long long *__restrict__ evals, float4 *s0, float4 *s1,
                                                        float4 *s2, uint32_t *wkeep) {
  __shared__ uint32_t bkeep[4][kWaves];
  uint32_t v, tx, ty;
  if (!raster_tile(views, rblock_start, nviews, total_rblocks, ro, b, v, tx, ty)) return;
  const uint32_t cap = stage_capacity(blend_min);
  uint32_t *lst = reinterpret_cast<uint32_t *>(s2 + cap);
  const ViewDev &vw = views[v];
  const uint32_t shift = vw.bin_shift;
  const uint32_t width = vw.fp.width, height = vw.fp.height;
  const uint32_t t = threadIdx.x, w = t / kWave, lane = __lane_id();
  const uint32_t x = tx * kTile + (w & 1u) * 8u + (lane & 7u), y = ty * kTile + (w >> 1) * 8u + (lane >> 3);
  const bool inside = x < width && y < height;
  const float pnx = (2.0f * ((float)x + 0.5f)) / (float)width - 1.0f;
  const float pny = 1.0f - (2.0f * ((float)y + 0.5f)) / (float)height;
  const uint2 rg = ranges[vw.ctile_off + (ty >> shift) * vw.fp.ctiles_x + (tx >> shift)];
  const uint32_t tx0 = tx * kTile, tx1 = tx0 + kTile - 1, ty0 = ty * kTile, ty1 = ty0 + kTile - 1;
  const float exlo = (2.0f * (float)tx0) / (float)width - 1.0f, exhi = (2.0f * (float)(tx1 + 1)) / (float)width - 1.0f;
  const float eyhi = 1.0f - (2.0f * (float)ty0) / (float)height, eylo = 1.0f - (2.0f * (float)(ty1 + 1)) / (float)height;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (kWave - lane));
  float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
  bool done = !inside;
  uint32_t passed = 0, tested = 0;
  uint32_t staged = 0;
  uint32_t bn[4] = {0, 0, 0, 0};  // entries of each block's list
  for (uint32_t start = rg.x;;) {
    if (__syncthreads_count(!done) == 0) break;
    const bool more = start < rg.y;
    if (more) {
      const uint32_t k = start + t;
      uint32_t s = 0, mask = 0;
      bool keep = false;
      float4 ra = make_float4(0.0f, 0.0f, 0.0f, 0.0f), rb = ra;
      if (k < rg.y) {
        s = list[k];
        const uint2 r = rect[s];
        const uint32_t rx0 = r.x & 0xFFFFu, ry0 = r.x >> 16, rx1 = r.y & 0xFFFFu, ry1 = r.y >> 16;
        keep = rx0 <= tx1 && rx1 >= tx0 && ry0 <= ty1 && ry1 >= ty0;
        if (keep) {  // drop splats whose ellipse misses the tile
          ra = rec[3 * (size_t)s + 0];
          rb = rec[3 * (size_t)s + 1];
          keep = ellipse_meets_rect(ra, rb, exlo, exhi, eylo, eyhi);
        }
        if (keep) {
          const bool left = rx0 <= tx0 + 7u, right = rx1 >= tx0 + 8u;
          const bool top = ry0 <= ty0 + 7u, bottom = ry1 >= ty0 + 8u;
          mask = (left && top ? 1u : 0u) | (right && top ? 2u : 0u) | (left && bottom ? 4u : 0u) |
                 (right && bottom ? 8u : 0u);
        }
      }
      const uint64_t bal = __ballot(keep);
      uint64_t bb[4];
#pragma unroll
      for (uint32_t q = 0; q < 4; ++q) bb[q] = __ballot((mask >> q) & 1u);
      if (lane == 0) {
        wkeep[w] = (uint32_t)__popcll(bal);
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) bkeep[q][w] = (uint32_t)__popcll(bb[q]);
      }
      __syncthreads();
      uint32_t off = staged, kept = 0;
#pragma unroll
      for (uint32_t q = 0; q < kWaves; ++q) {
        off += q < w ? wkeep[q] : 0u;
        kept += wkeep[q];
      }
      if (keep) {
        const uint32_t pos = off + (uint32_t)__popcll(bal & lt);
        s0[pos] = ra;
        s1[pos] = rb;
        s2[pos] = rec[3 * (size_t)s + 2];
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) {
          if ((mask >> q) & 1u) {
            uint32_t bo = bn[q];
#pragma unroll
            for (uint32_t p = 0; p < kWaves; ++p) bo += p < w ? bkeep[q][p] : 0u;
            lst[q * cap + bo + (uint32_t)__popcll(bb[q] & lt)] = pos;
          }
        }
      }
#pragma unroll
      for (uint32_t q = 0; q < 4; ++q)
#pragma unroll
        for (uint32_t p = 0; p < kWaves; ++p) bn[q] += bkeep[q][p];
      staged += kept;
      start += kThreads;
      __syncthreads();
    }
    if (staged >= blend_min || (!more || start >= rg.y)) {
      const uint32_t n = bn[w];
      if (n > 0) {
        constexpr uint32_t U = 4;
        const uint32_t *bl = lst + w * cap;
        bool active = !done;
        for (uint32_t j0 = 0; j0 < n; j0 += U) {
          if (__ballot(active) == 0) break;  // wave-uniform
          float4 ga[U], gb[U], gc[U];
#pragma unroll
          for (uint32_t u = 0; u < U; ++u) {
            const uint32_t pos = bl[min(j0 + u, n - 1)];
            ga[u] = s0[pos];
            gb[u] = s1[pos];
            gc[u] = s2[pos];
          }
#pragma unroll
          for (uint32_t u = 0; u < U; ++u) {
            const bool valid = j0 + u < n;
            const float4 a = ga[u];
            const float4 bq = gb[u];
            const float dx = pnx - a.x;
            const float dy = pny - a.y;
            const float lx = fmaf(a.w, dy, a.z * dx);
            const float ly = fmaf(bq.y, dy, bq.x * dx);
            const float r2 = fmaf(ly, ly, lx * lx);
            const bool hit = active && valid && r2 <= bq.z;
            if (COUNT) {
              tested += (active && valid) ? 1u : 0u;
              passed += hit ? 1u : 0u;
            }
            if (hit) {
              blend_hit(r2, bq.w, gc[u], T, cr, cg, cb);
              active = !(T < 1e-4f);
            }
          }
        }
        done = !active;
      }
      staged = 0;
#pragma unroll
      for (uint32_t q = 0; q < 4; ++q) bn[q] = 0;
      if (start >= rg.y) break;
    }
  }
  if (inside) write_pixel(vw, x, y, cr, cg, cb, T, img32, img8);
  if (COUNT) {
    uint32_t wp = passed, u = tested;
#pragma unroll
    for (uint32_t off = kWave / 2; off > 0; off >>= 1) {
      wp += __shfl_down(wp, off, kWave);
      u += __shfl_down(u, off, kWave);
    }
    if (lane == 0 && wp) atomicAdd(evals, (unsigned long long)wp);
    if (lane == 0 && u) atomicAdd(evals + 1, (unsigned long long)u);
  }
}

template <bool COUNT, int LOOP>
__device__ __forceinline__ void raster_workgroup(uint32_t b, const ViewDev *__restrict__ views,
                                                 const uint32_t *__restrict__ rblock_start, uint32_t nviews,
                                                 uint32_t total_rblocks, const RasterOrder &ro, uint32_t blend_min,
                                                 const float4 *__restrict__ rec, const uint2 *__restrict__ rect,
                                                 const uint2 *__restrict__ ranges, const uint32_t *__restrict__ list,
                                                 float4 *__restrict__ img32, uchar4 *__restrict__ img8,
                                                 unsigned long long *__restrict__ evals, float4 *s0, float4 *s1,
                                                 float4 *s2, uint32_t *wkeep) {
  if constexpr (LOOP == 2) {
    raster_workgroup_blocks<COUNT>(b, views, rblock_start, nviews, total_rblocks, ro, blend_min, rec, rect, ranges,
                                   list, img32, img8, evals, s0, s1, s2, wkeep);
    return;
  }
  uint32_t v, tx, ty;
  if (!raster_tile(views, rblock_start, nviews, total_rblocks, ro, b, v, tx, ty)) return;
  const ViewDev &vw = views[v];
  const uint32_t shift = vw.bin_shift;
  const uint32_t width = vw.fp.width, height = vw.fp.height;
  const uint32_t t = threadIdx.x, w = t / kWave, lane = __lane_id();
  const uint32_t x = tx * kTile + (t % kTile), y = ty * kTile + (t / kTile);
  const bool inside = x < width && y < height;
  const float pnx = (2.0f * ((float)x + 0.5f)) / (float)width - 1.0f;
  const float pny = 1.0f - (2.0f * ((float)y + 0.5f)) / (float)height;
  const uint2 rg = ranges[vw.ctile_off + (ty >> shift) * vw.fp.ctiles_x + (tx >> shift)];
  const uint32_t tx0 = tx * kTile, tx1 = tx0 + kTile - 1, ty0 = ty * kTile, ty1 = ty0 + kTile - 1;
  // the tile's pixel edges in NDC (x right, y up)
  const float exlo = (2.0f * (float)tx0) / (float)width - 1.0f, exhi = (2.0f * (float)(tx1 + 1)) / (float)width - 1.0f;
  const float eyhi = 1.0f - (2.0f * (float)ty0) / (float)height, eylo = 1.0f - (2.0f * (float)(ty1 + 1)) / (float)height;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (kWave - lane));
  float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
  bool done = !inside;
  uint32_t passed = 0, tested = 0;
  uint32_t staged = 0;
  for (uint32_t start = rg.x;;) {
    if (__syncthreads_count(!done) == 0) break;
    const bool more = start < rg.y;
    if (more) {
      const uint32_t k = start + t;
      uint32_t s = 0;
      bool keep = false;
      float4 ra = make_float4(0.0f, 0.0f, 0.0f, 0.0f), rb = ra;
      if (k < rg.y) {
        s = list[k];
        if (shift == 0) {
          keep = true;
        } else {
          const uint2 r = rect[s];
          keep = (r.x & 0xFFFFu) <= tx1 && (r.y & 0xFFFFu) >= tx0 && (r.x >> 16) <= ty1 && (r.y >> 16) >= ty0;
        }
        if (LOOP != 0 && keep) {  // drop splats whose ellipse misses the tile (the Linear loop keeps them)
          ra = rec[3 * (size_t)s + 0];
          rb = rec[3 * (size_t)s + 1];
          keep = ellipse_meets_rect(ra, rb, exlo, exhi, eylo, eyhi);
        }
      }
      const uint64_t bal = __ballot(keep);
      if (lane == 0) wkeep[w] = (uint32_t)__popcll(bal);
      __syncthreads();
      uint32_t off = staged, kept = 0;
#pragma unroll
      for (uint32_t q = 0; q < kWaves; ++q) {
        off += q < w ? wkeep[q] : 0u;
        kept += wkeep[q];
      }
      if (keep) {
        const uint32_t pos = off + (uint32_t)__popcll(bal & lt);
        s0[pos] = LOOP != 0 ? ra : rec[3 * (size_t)s + 0];
        s1[pos] = LOOP != 0 ? rb : rec[3 * (size_t)s + 1];
        s2[pos] = rec[3 * (size_t)s + 2];
      }
      staged += kept;
      start += kThreads;
      __syncthreads();
    }
    if (staged >= blend_min || (!more || start >= rg.y)) {
      if (staged > 0)
        blend_stage<LOOP, COUNT>(s0, s1, s2, staged, pnx, pny, T, cr, cg, cb, done, passed, tested);
      staged = 0;
      if (start >= rg.y) break;
    }
  }
  if (inside) write_pixel(vw, x, y, cr, cg, cb, T, img32, img8);
  if (COUNT) {
    uint32_t wp = passed, u = tested;
#pragma unroll
    for (uint32_t off = kWave / 2; off > 0; off >>= 1) {
      wp += __shfl_down(wp, off, kWave);
      u += __shfl_down(u, off, kWave);
    }
    if (lane == 0 && wp) atomicAdd(evals, (unsigned long long)wp);
    if (lane == 0 && u) atomicAdd(evals + 1, (unsigned long long)u);
  }
}

template <bool COUNT, int LOOP>
__global__ __launch_bounds__(kThreads) void k_raster(const ViewDev *__restrict__ views,
                                                     const uint32_t *__restrict__ rblock_start, uint32_t nviews,
                                                     uint32_t total_rblocks, RasterOrder ro, uint32_t blend_min,
                                                     const float4 *__restrict__ rec, const uint2 *__restrict__ rect,
                                                     const uint2 *__restrict__ ranges,
                                                     const uint32_t *__restrict__ list, float4 *__restrict__ img32,
                                                     uchar4 *__restrict__ img8,
                                                     unsigned long long *__restrict__ evals, Clears clears,
                                                     WordCopy copy) {
  raster_prologue(clears, copy);
  // the LDS stage, three arrays of stage_capacity(blend_min) records (sized
  // at launch: a smaller stage leaves room for more workgroups per CU)
  extern __shared__ float4 stage_lds[];
  const uint32_t cap = stage_capacity(blend_min);
  __shared__ uint32_t wkeep[kWaves];
  raster_workgroup<COUNT, LOOP>(blockIdx.x, views, rblock_start, nviews, total_rblocks, ro, blend_min, rec, rect,
                                ranges, list, img32, img8, evals, stage_lds, stage_lds + cap, stage_lds + 2 * cap,
                                wkeep);
}

// Persistent form of k_raster (the Unrolled loop): a grid of about what the
// chip holds at once takes the raster workgroups' work items from queues.
// Workgroup g runs on XCD x = g % kXcds and serves queue (x, y), y = (g /
// kXcds) % kQueueLanes, which hands out the items b = x + kXcds (y +
// kQueueLanes ticket), in order: the items keep the XCD placement and the
// order of the plain launch (the raster tiles of one binning tile on one XCD,
// heaviest lists first). Tickets are device-scope atomics, served where the
// XCDs meet, so the counters are spread over kXcds * kQueueLanes cache lines
// (kQueueStride words apart) and a workgroup asks for its next ticket before
// working on its current item. The counters must be zero at launch.
template <bool COUNT>
__global__ __launch_bounds__(kThreads) void k_raster_queue(const ViewDev *__restrict__ views,
                                                           const uint32_t *__restrict__ rblock_start, uint32_t nviews,
                                                           uint32_t total_rblocks, RasterOrder ro, uint32_t blend_min,
                                                           const float4 *__restrict__ rec,
                                                           const uint2 *__restrict__ rect,
                                                           const uint2 *__restrict__ ranges,
                                                           const uint32_t *__restrict__ list,
                                                           float4 *__restrict__ img32, uchar4 *__restrict__ img8,
                                                           unsigned long long *__restrict__ evals,
                                                           uint32_t *__restrict__ queue, Clears clears,
                                                           WordCopy copy) {
  raster_prologue(clears, copy);
  extern __shared__ float4 stage_lds[];
  const uint32_t cap = stage_capacity(blend_min);
  __shared__ uint32_t wkeep[kWaves];
  __shared__ uint32_t s_ticket;
  const uint32_t x = blockIdx.x % kXcds, y = (blockIdx.x / kXcds) % kQueueLanes;
  uint32_t *counter = queue + (size_t)(x * kQueueLanes + y) * kQueueStride;
  uint32_t next = 0;
  if (threadIdx.x == 0) next = atomicAdd(counter, 1u);
  for (;;) {
    if (threadIdx.x == 0) s_ticket = next;
    __syncthreads();
    const uint32_t b = x + kXcds * (y + kQueueLanes * s_ticket);
    __syncthreads();  // every thread has the item before s_ticket changes
    if (b >= total_rblocks) break;  // uniform: this queue is drained
    if (threadIdx.x == 0) next = atomicAdd(counter, 1u);  // the next ticket, fetched meanwhile
    raster_workgroup<COUNT, 1>(b, views, rblock_start, nviews, total_rblocks, ro, blend_min, rec, rect, ranges,
                               list, img32, img8, evals, stage_lds, stage_lds + cap, stage_lds + 2 * cap, wkeep);
    __syncthreads();  // the stage and wkeep are free for the next item
  }
}

// One pixel's state in the two-pixel blend.
struct PixelAcc {
  float T, cr, cg, cb;
  bool live;
  uint32_t passed, tested;
};

// Coverage test and blend of staged splat j for one pixel at (pnx, pny):
// the same operations, in the same order, as the one-pixel blend_stage.
template <bool COUNT>
__device__ __forceinline__ void blend_one(PixelAcc &p, float pnx, float pny, const float4 &a, const float4 &b,
                                          const float4 *s2, uint32_t j, bool real) {
  const float dx = pnx - a.x;
  const float dy = pny - a.y;
  const float lx = fmaf(a.w, dy, a.z * dx);
  const float ly = fmaf(b.y, dy, b.x * dx);
  const float r2 = fmaf(ly, ly, lx * lx);
  const bool hit = p.live && r2 <= b.z;
  if (COUNT) {
    p.tested += (p.live && real) ? 1u : 0u;
    p.passed += hit ? 1u : 0u;
  }
  if (hit) {
    blend_hit(r2, b.w, s2[j], p.T, p.cr, p.cg, p.cb);
    p.live = !(p.T < 1e-4f);
  }
}

// Blend of n staged splats for the two pixels of a lane: one LDS read of
// each splat's coverage terms serves both pixels. Wave-uniform loop, four
// splats per liveness check (padding entries cannot pass coverage).
template <bool COUNT>
__device__ __forceinline__ void blend_stage2(const float4 *s0, float4 *s1, const float4 *s2, uint32_t n, float pnx,
                                             float pny0, float pny1, PixelAcc &p0, PixelAcc &p1) {
  constexpr uint32_t U = 4;
  const uint32_t nu = (n + U - 1) & ~(U - 1);
  const uint32_t t = threadIdx.x;
  if (t < nu - n) s1[n + t] = make_float4(0.0f, 0.0f, -1.0f, 0.0f);
  __syncthreads();
  for (uint32_t j0 = 0; j0 < nu; j0 += U) {
    if (__ballot(p0.live || p1.live) == 0) break;  // wave-uniform
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      const uint32_t j = j0 + u;
      const float4 a = s0[j];
      const float4 b = s1[j];
      blend_one<COUNT>(p0, pnx, pny0, a, b, s2, j, j < n);
      blend_one<COUNT>(p1, pnx, pny1, a, b, s2, j, j < n);
    }
  }
}

// Two vertically adjacent 16x16 tiles (tile rows ty, ty + 1, ty even) of
// one binning tile per workgroup, two pixels per thread: wave w covers rows
// 8w .. 8w+7 of the 16x32 region, lane l the pixels (l % 16, 8w + 2(l/16))
// and the one below. The entries of the binning tile's list whose pixel
// bounds overlap the region are staged in order as in k_raster, and each
// pixel sees the same splats in the same order with the same arithmetic, so
// the images equal k_raster's bit for bit. Needs bin shift >= 1 (both tile
// rows in one binning tile); workgroups of odd tile rows exit at once.
template <bool COUNT>
__global__ __launch_bounds__(kThreads) void k_raster_rows2(const ViewDev *__restrict__ views,
                                                           const uint32_t *__restrict__ rblock_start, uint32_t nviews,
                                                           uint32_t total_rblocks, RasterOrder ro, uint32_t blend_min,
                                                           const float4 *__restrict__ rec,
                                                           const uint2 *__restrict__ rect,
                                                           const uint2 *__restrict__ ranges,
                                                           const uint32_t *__restrict__ list,
                                                           float4 *__restrict__ img32, uchar4 *__restrict__ img8,
                                                           unsigned long long *__restrict__ evals, Clears clears,
                                                           WordCopy copy) {
  raster_prologue(clears, copy);
  extern __shared__ float4 stage_lds[];
  const uint32_t cap = stage_capacity(blend_min);
  float4 *s0 = stage_lds, *s1 = stage_lds + cap, *s2 = stage_lds + 2 * cap;
  __shared__ uint32_t wkeep[kWaves];
  uint32_t v, tx, ty;
  if (!raster_tile(views, rblock_start, nviews, total_rblocks, ro, blockIdx.x, v, tx, ty)) return;
  const ViewDev &vw = views[v];
  const uint32_t shift = vw.bin_shift;
  const uint32_t width = vw.fp.width, height = vw.fp.height;
  if (ty & 1u) return;  // covered by the workgroup of row ty - 1
  const uint32_t t = threadIdx.x, w = t / kWave, lane = __lane_id();
  const uint32_t x = tx * kTile + (lane % kTile);
  const uint32_t y0 = ty * kTile + 8 * w + 2 * (lane / kTile), y1 = y0 + 1;
  const bool in0 = x < width && y0 < height, in1 = x < width && y1 < height;
  const float pnx = (2.0f * ((float)x + 0.5f)) / (float)width - 1.0f;
  const float pny0 = 1.0f - (2.0f * ((float)y0 + 0.5f)) / (float)height;
  const float pny1 = 1.0f - (2.0f * ((float)y1 + 0.5f)) / (float)height;
  const uint2 rg = ranges[vw.ctile_off + (ty >> shift) * vw.fp.ctiles_x + (tx >> shift)];
  const uint32_t tx0 = tx * kTile, tx1 = tx0 + kTile - 1, ty0 = ty * kTile, ty1 = ty0 + 2 * kTile - 1;
  const float exlo = (2.0f * (float)tx0) / (float)width - 1.0f, exhi = (2.0f * (float)(tx1 + 1)) / (float)width - 1.0f;
  const float eyhi = 1.0f - (2.0f * (float)ty0) / (float)height, eylo = 1.0f - (2.0f * (float)(ty1 + 1)) / (float)height;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (kWave - lane));
  PixelAcc p0{1.0f, 0.0f, 0.0f, 0.0f, in0, 0u, 0u}, p1{1.0f, 0.0f, 0.0f, 0.0f, in1, 0u, 0u};
  uint32_t staged = 0;
  for (uint32_t start = rg.x;;) {
    if (__syncthreads_count(p0.live || p1.live) == 0) break;
    const bool more = start < rg.y;
    if (more) {
      const uint32_t k = start + t;
      uint32_t s = 0;
      bool keep = false;
      float4 ra = make_float4(0.0f, 0.0f, 0.0f, 0.0f), rb = ra;
      if (k < rg.y) {
        s = list[k];
        const uint2 r = rect[s];
        keep = (r.x & 0xFFFFu) <= tx1 && (r.y & 0xFFFFu) >= tx0 && (r.x >> 16) <= ty1 && (r.y >> 16) >= ty0;
        if (keep) {
          ra = rec[3 * (size_t)s + 0];
          rb = rec[3 * (size_t)s + 1];
          keep = ellipse_meets_rect(ra, rb, exlo, exhi, eylo, eyhi);
        }
      }
      const uint64_t bal = __ballot(keep);
      if (lane == 0) wkeep[w] = (uint32_t)__popcll(bal);
      __syncthreads();
      uint32_t off = staged, kept = 0;
#pragma unroll
      for (uint32_t q = 0; q < kWaves; ++q) {
        off += q < w ? wkeep[q] : 0u;
        kept += wkeep[q];
      }
      if (keep) {
        const uint32_t pos = off + (uint32_t)__popcll(bal & lt);
        s0[pos] = ra;
        s1[pos] = rb;
        s2[pos] = rec[3 * (size_t)s + 2];
      }
      staged += kept;
      start += kThreads;
      __syncthreads();
    }
    if (staged >= blend_min || (!more || start >= rg.y)) {
      if (staged > 0) blend_stage2<COUNT>(s0, s1, s2, staged, pnx, pny0, pny1, p0, p1);
      staged = 0;
      if (start >= rg.y) break;
    }
  }
  if (in0) write_pixel(vw, x, y0, p0.cr, p0.cg, p0.cb, p0.T, img32, img8);
  if (in1) write_pixel(vw, x, y1, p1.cr, p1.cg, p1.cb, p1.T, img32, img8);
  if (COUNT) {
    uint32_t wp = p0.passed + p1.passed, u = p0.tested + p1.tested;
#pragma unroll
    for (uint32_t off = kWave / 2; off > 0; off >>= 1) {
      wp += __shfl_down(wp, off, kWave);
      u += __shfl_down(u, off, kWave);
    }
    if (lane == 0 && wp) atomicAdd(evals, (unsigned long long)wp);
    if (lane == 0 && u) atomicAdd(evals + 1, (unsigned long long)u);
  }
}

uint32_t div_up(size_t a, size_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

FrameParams make_frame_params(const float position[3], const float rotation_xyzw[4], float fov, float near_plane,
                              float far_plane, uint32_t width, uint32_t height) {
  FrameParams p{};
  for (int a = 0; a < 3; ++a) p.cam[a] = position[a];
  view_rows(rotation_xyzw, p.rows);
  p.near_plane = near_plane;
  p.far_plane = far_plane;
  // make_surface_render_params (lib.rs:889-920) and project_splat's
  // constants (wgsl:143-185); pow(x, 2.0) evaluated as x*x.
  p.aspect = std::max((float)width / (float)std::max(height, 1u), 1e-6f);
  const float tan_half = std::tan(fov * 0.5f);
  p.f = 1.0f / tan_half;
  p.fx = p.f / p.aspect;
  p.fy = p.f;
  p.lim_x = 1.3f * (tan_half * p.aspect);
  p.lim_y = 1.3f * tan_half;
  const float bx = 0.3f * (2.0f / std::max((float)width, 1.0f));
  const float by = 0.3f * (2.0f / std::max((float)height, 1.0f));
  p.blur_x = bx * bx;
  p.blur_y = by * by;
  p.width = width;
  p.height = height;
  p.tiles_x = (width + kTile - 1) / kTile;
  p.tiles_y = (height + kTile - 1) / kTile;
  p.ctiles_x = p.tiles_x;
  p.ctiles_y = p.tiles_y;
  return p;
}

void set_bin_tiles(FrameParams &fp, uint32_t shift) {
  fp.ctiles_x = (fp.tiles_x + (1u << shift) - 1) >> shift;
  fp.ctiles_y = (fp.tiles_y + (1u << shift) - 1) >> shift;
}

namespace {

// Calls f(std::integral_constant<uint32_t, F>) for the scene format F = fmt.
template <class F>
void with_format(uint32_t fmt, F &&f) {
  switch (fmt) {
    case 0: f(std::integral_constant<uint32_t, 0>()); break;
    case 1: f(std::integral_constant<uint32_t, 1>()); break;
    case 2: f(std::integral_constant<uint32_t, 2>()); break;
    case 3: f(std::integral_constant<uint32_t, 3>()); break;
    case 4: f(std::integral_constant<uint32_t, 4>()); break;
    case kPackedFormat + 0: f(std::integral_constant<uint32_t, kPackedFormat + 0>()); break;
    case kPackedFormat + 1: f(std::integral_constant<uint32_t, kPackedFormat + 1>()); break;
    case kPackedFormat + 2: f(std::integral_constant<uint32_t, kPackedFormat + 2>()); break;
    case kPackedFormat + 3: f(std::integral_constant<uint32_t, kPackedFormat + 3>()); break;
    default: throw Error(Code::Unsupported, "unknown scene format " + std::to_string(fmt));
  }
}

}  // namespace

void launch_project(uint32_t fmt, const ViewDev *views, const uint32_t *seg_start, uint32_t nviews,
                    uint32_t total_slots, uint32_t *keys, uint32_t *vals, float4 *rec, uint2 *rect, const Clears &clr,
                    hipStream_t st, uint2 *key_range, const InlineTable *table) {
  if (total_slots == 0 || nviews == 0) return;
  if (table && (nviews != 1 || table->count > kInlineTableWords || !table->dst ||
                table->view_word != kInlineViewWord || kInlineViewWord * 4 + sizeof(ViewDev) > table->count * 4))
    throw Error(Code::Internal, "inline view table of one view expected");
  const dim3 grid(div_up(total_slots, kThreads)), block(kThreads);
  with_format(fmt, [&](auto F) {
    if (table)
      k_project<decltype(F)::value, true><<<grid, block, 0, st>>>(views, seg_start, nviews, keys, vals, rec, rect,
                                                                  clr, key_range, *table);
    else
      k_project<decltype(F)::value, false><<<grid, block, 0, st>>>(views, seg_start, nviews, keys, vals, rec, rect,
                                                                   clr, key_range, InlineTable());
  });
  GS_HIP(hipGetLastError());
}

void launch_project_groups(uint32_t fmt, const ViewDev *views, const ProjGroup *groups, uint32_t ngroups,
                           uint32_t chunk_lo, uint32_t chunks, uint32_t *keys, uint32_t *vals, float4 *rec,
                           uint2 *rect, const Clears &clr, hipStream_t st) {
  if (chunks == 0 || ngroups == 0) return;
  const dim3 grid(chunks), block(kThreads);
  with_format(fmt, [&](auto F) {
    k_project_groups<decltype(F)::value>
        <<<grid, block, 0, st>>>(views, groups, ngroups, chunk_lo, keys, vals, rec, rect, clr);
  });
  GS_HIP(hipGetLastError());
}

void launch_project_export(uint32_t fmt, const ViewDev *views, uint32_t view, uint32_t n, uint32_t *keys_out,
                           float *export12, hipStream_t st) {
  if (n == 0) return;
  const dim3 grid(div_up(n, kThreads)), block(kThreads);
  with_format(fmt, [&](auto F) {
    k_project_export<decltype(F)::value><<<grid, block, 0, st>>>(views, view, keys_out, export12);
  });
  GS_HIP(hipGetLastError());
}

void launch_tile_counts(const ViewDev *views, const uint32_t *seg_start, uint32_t nviews, uint32_t total_slots,
                        const uint32_t *skeys, const uint32_t *svals, const uint2 *rect,
                        uint32_t *counts_rev, uint32_t *offsets_rev, uint32_t *visible, uint32_t *block_sums,
                        uint32_t *block_pairs, uint32_t *rect_lo, uint32_t *rect_hi, const Clears &clr,
                        hipStream_t st) {
  if (total_slots % kCountBlock) throw Error(Code::Internal, "slot count not a multiple of the count block");
  if (total_slots == 0) return;
  const uint32_t grid = (total_slots / kCountBlock + kXcds - 1) / kXcds * kXcds;
  k_tile_counts<<<grid, kThreads, 0, st>>>(views, seg_start, nviews, total_slots, skeys,
                                                                  svals, rect, counts_rev, offsets_rev, visible,
                                                                  block_sums, block_pairs, rect_lo, rect_hi, clr);
  GS_HIP(hipGetLastError());
}

void launch_block_scan(uint32_t *block_sums, uint32_t blocks, uint32_t capacity, uint32_t *m_out, uint32_t *info,
                       const uint32_t *block_pairs, unsigned long long *raster_pairs, hipStream_t st) {
  k_block_scan<<<1, kBlockScanThreads, 0, st>>>(block_sums, blocks, capacity, m_out, info, block_pairs, raster_pairs);
  GS_HIP(hipGetLastError());
}

void launch_tile_emit(const ViewDev *views, const uint32_t *seg_start, uint32_t nviews, uint32_t total_slots,
                      const uint32_t *svals, const uint32_t *rect_lo, const uint32_t *rect_hi,
                      const uint32_t *counts_rev,
                      const uint32_t *offsets_rev, const uint32_t *block_sums, uint32_t capacity,
                      uint32_t *pair_keys, uint32_t *pair_vals, hipStream_t st, const PairHist &ph,
                      const EmitTotals &et) {
  if (total_slots == 0) return;
  if (ph.hist && (ph.passes > kPairHistPasses || ph.digit_bits > 9))
    throw Error(Code::Internal, "pair histogram of more than two passes or 9-bit digits");
  if (et.blocks && (et.blocks != total_slots / kCountBlock || et.blocks > kEmitScanMaxBlocks || !et.block_pairs ||
                    !et.m_out || !et.info || !et.raster_pairs))
    throw Error(Code::Internal, "emission block totals do not match the slots");
  k_tile_emit<<<total_slots / kThreads, kThreads, 0, st>>>(views, seg_start, nviews, total_slots, svals,
                                                             rect_lo, rect_hi, counts_rev, offsets_rev, block_sums,
                                                             capacity,
                                                             pair_keys, pair_vals, ph, et);
  GS_HIP(hipGetLastError());
}

void launch_raster_tile_counts(const ViewDev *views, uint32_t v, const ViewDev &host_view, const uint32_t *skeys,
                               const uint32_t *svals, const uint2 *rect, uint32_t *counts, hipStream_t st) {
  launch_fill_u32(counts, 0u, host_view.tiles, st);
  if (host_view.seg_len == 0) return;
  k_raster_tile_counts<<<div_up(host_view.seg_len, kThreads), kThreads, 0, st>>>(views, v, skeys, svals, rect,
                                                                                  counts);
  GS_HIP(hipGetLastError());
}

void launch_fill_u32(uint32_t *p, uint32_t value, size_t n, hipStream_t st) {
  if (n == 0) return;
  k_fill_u32<<<div_up(n, kThreads), kThreads, 0, st>>>(p, value, n);
  GS_HIP(hipGetLastError());
}

void launch_tile_ranges(const ViewDev *views, const uint32_t *pair_keys, const uint32_t *m_dev, uint32_t capacity,
                        uint2 *ranges, const Clears &clr, const WordCopy &copy, hipStream_t st) {
  const uint32_t grid = std::max<uint32_t>(std::min<uint32_t>(div_up(capacity, kThreads), kMaxGrid), 1);
  k_tile_ranges<<<grid, kThreads, 0, st>>>(views, pair_keys, m_dev, ranges, clr, copy);
  GS_HIP(hipGetLastError());
}

void launch_raster_order(const uint2 *ranges, uint32_t total_ctiles, uint32_t *counts, uint32_t *order,
                         hipStream_t st) {
  if (total_ctiles == 0) return;
  k_raster_order<<<div_up(total_ctiles, kThreads), kThreads, 0, st>>>(ranges, total_ctiles, counts, order);
  GS_HIP(hipGetLastError());
}

// Workgroups of the persistent raster: eight per CU (the raster's registers
// allow seven resident at once; the eighth of each CU starts as soon as one
// ends and finds its queue's remaining items or none), in whole runs of
// kXcds * kQueueLanes so that every queue has workgroups.
static uint32_t queue_grid() {
  static thread_local uint32_t cached = 0;
  if (cached) return cached;
  int dev = 0, cus = 0;
  GS_HIP(hipGetDevice(&dev));
  GS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  constexpr uint32_t run = kXcds * kQueueLanes;
  cached = std::max<uint32_t>(run, (uint32_t)std::max(cus, 1) * 8u / run * run);
  return cached;
}

void launch_raster(RasterVariant variant, const ViewDev *views, const uint32_t *rblock_start, uint32_t nviews,
                   uint32_t total_rblocks, const RasterOrder &ro, uint32_t blend_min, const float4 *rec, const uint2 *rect, const uint2 *ranges,
                   const uint32_t *list, float4 *img32, uchar4 *img8, hipStream_t st, unsigned long long *evals,
                   uint32_t *queue, const Clears &clears, const WordCopy &copy) {
  if (total_rblocks == 0 || nviews == 0) return;
  if (blend_min == 0 || blend_min > kThreads) throw Error(Code::InvalidArgument, "blend threshold outside 1..256");
  const dim3 grid(total_rblocks), block(kThreads);
  const size_t lds = 3 * (size_t)stage_capacity(blend_min) * sizeof(float4);
  if (variant == RasterVariant::Blocks8) {  // the stage plus four position lists of stage_capacity words
    const size_t lds8 = lds + (size_t)stage_capacity(blend_min) * sizeof(float4);
    if (evals)
      k_raster<true, 2><<<grid, block, lds8, st>>>(views, rblock_start, nviews, total_rblocks, ro, blend_min, rec, rect,
                                                  ranges, list, img32, img8, evals, clears, copy);
    else
      k_raster<false, 2><<<grid, block, lds8, st>>>(views, rblock_start, nviews, total_rblocks, ro, blend_min, rec,
                                                   rect, ranges, list, img32, img8, evals, clears, copy);
    GS_HIP(hipGetLastError());
    return;
  }
  if (variant == RasterVariant::Queue) {
    if (!queue) throw Error(Code::Internal, "persistent raster without its queue words");
    constexpr uint32_t run = kXcds * kQueueLanes;  // every queue keeps at least one workgroup
    const dim3 qgrid(std::min(queue_grid(), (total_rblocks + run - 1) / run * run));
    if (evals)
      k_raster_queue<true><<<qgrid, block, lds, st>>>(views, rblock_start, nviews, total_rblocks, ro, blend_min, rec,
                                                      rect, ranges, list, img32, img8, evals, queue, clears, copy);
    else
      k_raster_queue<false><<<qgrid, block, lds, st>>>(views, rblock_start, nviews, total_rblocks, ro, blend_min, rec,
                                                       rect, ranges, list, img32, img8, evals, queue, clears, copy);
    GS_HIP(hipGetLastError());
    return;
  }
  if (variant == RasterVariant::Rows2) {  // every view at bin shift >= 1 (the caller's duty)
    if (evals)
      k_raster_rows2<true><<<grid, block, lds, st>>>(views, rblock_start, nviews, total_rblocks, ro, blend_min, rec, rect, ranges,
                                                   list, img32, img8, evals, clears, copy);
    else
      k_raster_rows2<false><<<grid, block, lds, st>>>(views, rblock_start, nviews, total_rblocks, ro, blend_min, rec, rect, ranges,
                                                    list, img32, img8, evals, clears, copy);
    GS_HIP(hipGetLastError());
    return;
  }
  const bool unrolled = variant != RasterVariant::Linear;
  if (unrolled && evals)
    k_raster<true, 1><<<grid, block, lds, st>>>(views, rblock_start, nviews, total_rblocks, ro, blend_min, rec, rect, ranges, list,
                                              img32, img8, evals, clears, copy);
  else if (unrolled)
    k_raster<false, 1><<<grid, block, lds, st>>>(views, rblock_start, nviews, total_rblocks, ro, blend_min, rec, rect, ranges,
                                               list, img32, img8, evals, clears, copy);
  else if (evals)
    k_raster<true, 0><<<grid, block, lds, st>>>(views, rblock_start, nviews, total_rblocks, ro, blend_min, rec, rect, ranges, list,
                                              img32, img8, evals, clears, copy);
  else
    k_raster<false, 0><<<grid, block, lds, st>>>(views, rblock_start, nviews, total_rblocks, ro, blend_min, rec, rect, ranges,
                                               list, img32, img8, evals, clears, copy);
  GS_HIP(hipGetLastError());
}

}  // namespace gs
