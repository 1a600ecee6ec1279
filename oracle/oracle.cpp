// oracle.cpp -- CPU restatement of the gsplat-rs SortedAlpha render path.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the HIP path,
// never linked into or called by the product library.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off, no fast-math). Paths in
// citations are relative to the gsplat-rs repository root.

#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

// Planted mistakes for the pinning tests' self-check: tests/
// test_oracle_mutations.py builds this file with -DORACLE_MUTATION=k into a
// temporary directory and expects the pins of tests/test_oracle_projection.py
// to fail. Every normal build has ORACLE_MUTATION 0 (no mistake).
//   1 sign of j02 flipped            5 NDC-box cull on |axis_u.x| only
//   2 aspect dropped from lim_x      6 blur term left out
//   3 64x anisotropy cap removed     7 tile rect from the quad extent
//   4 upper radius clamp removed     8 lower radius clamp removed
#ifndef ORACLE_MUTATION
#define ORACLE_MUTATION 0
#endif

namespace {

struct Q {
  float x, y, z, w;
};

// quat_normalize, crates/gsplat-render-wgpu/src/lib.rs:1380-1387.
Q quat_normalize(Q q) {
  float norm2 = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
  if (norm2 <= 0.0f) return Q{0.0f, 0.0f, 0.0f, 1.0f};
  float inv = 1.0f / std::sqrt(norm2);
  return Q{q.x * inv, q.y * inv, q.z * inv, q.w * inv};
}

// quat_inverse, lib.rs:1374-1378.
Q quat_inverse(Q q) {
  q = quat_normalize(q);
  return Q{-q.x, -q.y, -q.z, q.w};
}

// quat_to_mat3, lib.rs:1389-1407 (normalises its input again).
void quat_to_mat3(Q q, float m[3][3]) {
  q = quat_normalize(q);
  float x = q.x, y = q.y, z = q.z, w = q.w;
  float xx = x * x, yy = y * y, zz = z * z;
  float xy = x * y, xz = x * z, yz = y * z;
  float wx = w * x, wy = w * y, wz = w * z;
  m[0][0] = 1.0f - 2.0f * (yy + zz);
  m[0][1] = 2.0f * (xy - wz);
  m[0][2] = 2.0f * (xz + wy);
  m[1][0] = 2.0f * (xy + wz);
  m[1][1] = 1.0f - 2.0f * (xx + zz);
  m[1][2] = 2.0f * (yz - wx);
  m[2][0] = 2.0f * (xz - wy);
  m[2][1] = 2.0f * (yz + wx);
  m[2][2] = 1.0f - 2.0f * (xx + yy);
}

// mat3_mul, lib.rs:1409-1417 (row . column, full sums including zeros).
void mat3_mul(const float a[3][3], const float b[3][3], float out[3][3]) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) out[r][c] = a[r][0] * b[0][c] + a[r][1] * b[1][c] + a[r][2] * b[2][c];
}

struct View {
  float pos[3];
  float rows[3][3];
  float near_plane, far_plane, fov;
};

View make_view(const OracleCamera *cam) {
  View v;
  for (int i = 0; i < 3; ++i) v.pos[i] = cam->position[i];
  Q q{cam->rotation_xyzw[0], cam->rotation_xyzw[1], cam->rotation_xyzw[2], cam->rotation_xyzw[3]};
  quat_to_mat3(quat_inverse(q), v.rows);  // make_surface_render_params, lib.rs:896-897
  v.near_plane = cam->near_plane;
  v.far_plane = cam->far_plane;
  v.fov = cam->vertical_fov_radians;
  return v;
}

// world_to_camera_depth_with_view_row, lib.rs:1346-1354: left to right, no FMA.
float depth_of(const View &v, const float *p) {
  return v.rows[2][0] * (p[0] - v.pos[0]) + v.rows[2][1] * (p[1] - v.pos[1]) +
         v.rows[2][2] * (p[2] - v.pos[2]);
}

uint32_t f32_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

float bits_f32(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// 8-bit stable descending LSD radix over the key half of key<<32 | !value,
// sort/src/radix.rs:21-58 + :93-126 and sort/src/lib.rs:36-61, 104-107.
void radix_sort_key_bits(std::vector<uint64_t> &values) {
  const size_t n = values.size();
  std::vector<uint64_t> scratch(n);
  size_t counts[256];
  bool values_to_scratch = true;
  for (int shift = 32; shift < 64; shift += 8) {
    const std::vector<uint64_t> &in = values_to_scratch ? values : scratch;
    std::vector<uint64_t> &out = values_to_scratch ? scratch : values;
    std::fill(counts, counts + 256, size_t(0));  // count_histograms_scalar :93-98
    for (uint64_t v : in) counts[(v >> shift) & 0xFF] += 1;
    size_t offset = 0;  // descending_prefix_offsets :110-117
    for (int d = 255; d >= 0; --d) {
      size_t len = counts[d];
      counts[d] = offset;
      offset += len;
    }
    for (uint64_t v : in) out[counts[(v >> shift) & 0xFF]++] = v;  // scatter :119-126
    values_to_scratch = !values_to_scratch;
  }
  if (!values_to_scratch) values = scratch;
}

struct Proj {
  bool valid;
  float cx, cy, ux, uy, vx, vy, alpha;
};

struct FrameConsts {
  float aspect, f, fx, fy, lim_x, lim_y, blur_x, blur_y;
};

// Per-frame constants of project_splat (wgsl:143-185) with the params of
// make_surface_render_params (lib.rs:889-920). pow(x, 2.0) is stated as x*x
// (the reference CPU oracle's powi(2), lib.rs:1243-1244).
FrameConsts frame_consts(const View &v, uint32_t width, uint32_t height) {
  FrameConsts c;
  float w = (float)width;
  float h = (float)std::max(height, 1u);
  c.aspect = std::max(w / h, 1e-6f);
  float tan_half = std::tan(v.fov * 0.5f);
  c.f = 1.0f / tan_half;
  c.fx = c.f / c.aspect;
  c.fy = c.f;
  c.lim_x = 1.3f * (tan_half * c.aspect);
#if ORACLE_MUTATION == 2
  c.lim_x = 1.3f * tan_half;
#endif
  c.lim_y = 1.3f * tan_half;
  float px_x = 2.0f / std::max((float)width, 1.0f);
  float px_y = 2.0f / std::max((float)height, 1.0f);
  float bx = 0.3f * px_x, by = 0.3f * px_y;
  c.blur_x = bx * bx;
  c.blur_y = by * by;
#if ORACLE_MUTATION == 6
  c.blur_x = c.blur_y = 0.0f;
#endif
  return c;
}

// covariance_quadratic / covariance_bilinear, wgsl:95-109.
float cov_quadratic(const float c[6], const float r[3]) {
  return r[0] * r[0] * c[0] + 2.0f * r[0] * r[1] * c[1] + 2.0f * r[0] * r[2] * c[2] +
         r[1] * r[1] * c[3] + 2.0f * r[1] * r[2] * c[4] + r[2] * r[2] * c[5];
}
float cov_bilinear(const float c[6], const float a[3], const float b[3]) {
  float bx = c[0] * b[0] + c[1] * b[1] + c[2] * b[2];
  float by = c[1] * b[0] + c[3] * b[1] + c[4] * b[2];
  float bz = c[2] * b[0] + c[4] * b[1] + c[5] * b[2];
  return a[0] * bx + a[1] * by + a[2] * bz;
}

// project_splat, splat_surface_direct.wgsl:131-217.
Proj project_splat(const View &v, const FrameConsts &k, const float *p, const float cov[6],
                   float alpha) {
  Proj out{false, 2.0f, 2.0f, 0, 0, 0, 0, 0.0f};
  float rel[3] = {p[0] - v.pos[0], p[1] - v.pos[1], p[2] - v.pos[2]};
  const float *r0 = v.rows[0], *r1 = v.rows[1], *r2 = v.rows[2];
  float x = r0[0] * rel[0] + r0[1] * rel[1] + r0[2] * rel[2];
  float y = r1[0] * rel[0] + r1[1] * rel[1] + r1[2] * rel[2];
  float z = r2[0] * rel[0] + r2[1] * rel[1] + r2[2] * rel[2];
  if (z < v.near_plane || z > v.far_plane || z <= 1e-6f) return out;

  float inv_z = 1.0f / z;
  float x_ndc = ((x * k.f) * inv_z) / k.aspect;
  float y_ndc = (y * k.f) * inv_z;

  float cc[6] = {cov_quadratic(cov, r0),   cov_bilinear(cov, r0, r1), cov_bilinear(cov, r0, r2),
                 cov_quadratic(cov, r1),   cov_bilinear(cov, r1, r2), cov_quadratic(cov, r2)};
  // cc = (xx, xy, xz, yy, yz, zz) in camera space.

  float x_clamped = std::min(std::max(x / z, -k.lim_x), k.lim_x) * z;
  float y_clamped = std::min(std::max(y / z, -k.lim_y), k.lim_y) * z;
  float inv_z2 = inv_z * inv_z;
  float j00 = k.fx * inv_z;
  float j02 = -k.fx * x_clamped * inv_z2;
#if ORACLE_MUTATION == 1
  j02 = -j02;
#endif
  float j11 = k.fy * inv_z;
  float j12 = -k.fy * y_clamped * inv_z2;

  float cov01 = j00 * j11 * cc[1] + j00 * j12 * cc[2] + j02 * j11 * cc[4] + j02 * j12 * cc[5];
  float a = j00 * j00 * cc[0] + 2.0f * j00 * j02 * cc[2] + j02 * j02 * cc[5];
  float b = cov01;
  float c = j11 * j11 * cc[3] + 2.0f * j11 * j12 * cc[4] + j12 * j12 * cc[5];
  a = a + k.blur_x;
  c = c + k.blur_y;

  float apco2 = (a + c) * 0.5f;
  float amco2 = (a - c) * 0.5f;
  float term = std::sqrt(std::max(amco2 * amco2 + b * b, 0.0f));
  float major = std::max(apco2 + term, 1e-10f);
  float minor = std::max(apco2 - term, 1e-10f);

  float ud_x = 1.0f, ud_y = 0.0f;
  if (std::fabs(b) > 1e-8f) {
    // normalize2_or_default (wgsl:79-85); inverseSqrt stated as 1/sqrt.
    float vx = b, vy = major - a;
    float len2 = vx * vx + vy * vy;
    if (len2 > 1e-20f) {
      float inv = 1.0f / std::sqrt(len2);
      ud_x = vx * inv;
      ud_y = vy * inv;
    }
  } else if (a < c) {
    ud_x = 0.0f;
    ud_y = 1.0f;
  }
  float vd_x = -ud_y, vd_y = ud_x;

  float major_r = std::min(std::max(std::sqrt(major) * 3.0f, 1e-4f), 2.0f);
  float minor_r = std::min(std::max(std::sqrt(minor) * 3.0f, 1e-4f), 2.0f);
#if ORACLE_MUTATION == 4
  major_r = std::max(std::sqrt(major) * 3.0f, 1e-4f);
  minor_r = std::max(std::sqrt(minor) * 3.0f, 1e-4f);
#elif ORACLE_MUTATION == 8
  major_r = std::min(std::sqrt(major) * 3.0f, 2.0f);
  minor_r = std::min(std::sqrt(minor) * 3.0f, 2.0f);
#endif
#if ORACLE_MUTATION != 3
  major_r = std::min(major_r, minor_r * 64.0f);
#endif
  float ux = ud_x * major_r, uy = ud_y * major_r;
  float vx = vd_x * minor_r, vy = vd_y * minor_r;

  float extent_x = std::fabs(ux) + std::fabs(vx);
  float extent_y = std::fabs(uy) + std::fabs(vy);
#if ORACLE_MUTATION == 5
  extent_x = std::fabs(ux);
#endif
  if (x_ndc + extent_x < -1.0f || x_ndc - extent_x > 1.0f || y_ndc + extent_y < -1.0f ||
      y_ndc - extent_y > 1.0f)
    return out;

  out.valid = true;
  out.cx = x_ndc;
  out.cy = y_ndc;
  out.ux = ux;
  out.uy = uy;
  out.vx = vx;
  out.vy = vy;
  out.alpha = alpha;
  return out;
}

// eval_sh_rgb + eval_color, splat_surface_direct.wgsl:222-334. Accumulation
// order is the shader's left-to-right vec3 expression chain.
void eval_color(const OracleScene *s, uint32_t idx, const View &v, float alpha, float out_rgba[4]) {
  const float C0 = 0.28209479177387814f;
  const float C1 = 0.4886025119029199f;
  const float C2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                       -1.0925484305920792f, 0.5462742152960396f};
  const float C3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
                       0.3731763325901154f,  -0.4570457994644658f, 1.445305721320277f,
                       -0.5900435899266435f};
  const float C4[9] = {2.5033429417967046f,  -1.7701307697799304f, 0.9461746957575601f,
                       -0.6690465435572892f, 0.10578554691520431f, -0.6690465435572892f,
                       0.47308734787878004f, -1.7701307697799304f, 0.6258357354491761f};
  const float *p = s->positions + 3 * (size_t)idx;
  float d[3] = {p[0] - v.pos[0], p[1] - v.pos[1], p[2] - v.pos[2]};
  float len2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  float dir[3] = {0.0f, 0.0f, 1.0f};
  if (len2 > 1e-20f) {  // normalize3_or_default, wgsl:87-93 (inverseSqrt as 1/sqrt)
    float inv = 1.0f / std::sqrt(len2);
    dir[0] = d[0] * inv;
    dir[1] = d[1] * inv;
    dir[2] = d[2] * inv;
  }
  const float *dc = s->color_dc + 3 * (size_t)idx;
  float res[3] = {C0 * dc[0], C0 * dc[1], C0 * dc[2]};
  uint32_t degree = std::min(s->sh_degree, 4u);
  if (degree > 0 && s->sh_rest != nullptr) {
    uint32_t per_channel = (degree + 1) * (degree + 1) - 1;
    const float *base = s->sh_rest + (size_t)idx * per_channel * 3;
    float x = dir[0], y = dir[1], z = dir[2];
    auto rest = [&](int ch, uint32_t coeff) { return base[ch * per_channel + coeff]; };
    for (int ch = 0; ch < 3; ++ch) {
      float r = res[ch];
      r = r - C1 * y * rest(ch, 0) + C1 * z * rest(ch, 1) - C1 * x * rest(ch, 2);
      if (degree >= 2) {
        float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
        r = r + C2[0] * xy * rest(ch, 3) + C2[1] * yz * rest(ch, 4) +
            C2[2] * (2.0f * zz - xx - yy) * rest(ch, 5) + C2[3] * xz * rest(ch, 6) +
            C2[4] * (xx - yy) * rest(ch, 7);
        if (degree >= 3) {
          r = r + C3[0] * y * (3.0f * xx - yy) * rest(ch, 8) + C3[1] * xy * z * rest(ch, 9) +
              C3[2] * y * (4.0f * zz - xx - yy) * rest(ch, 10) +
              C3[3] * z * (2.0f * zz - 3.0f * xx - 3.0f * yy) * rest(ch, 11) +
              C3[4] * x * (4.0f * zz - xx - yy) * rest(ch, 12) + C3[5] * z * (xx - yy) * rest(ch, 13) +
              C3[6] * x * (xx - 3.0f * yy) * rest(ch, 14);
          if (degree >= 4) {
            r = r + C4[0] * xy * (xx - yy) * rest(ch, 15) + C4[1] * yz * (3.0f * xx - yy) * rest(ch, 16) +
                C4[2] * xy * (7.0f * zz - 1.0f) * rest(ch, 17) +
                C4[3] * yz * (7.0f * zz - 3.0f) * rest(ch, 18) +
                C4[4] * (zz * (35.0f * zz - 30.0f) + 3.0f) * rest(ch, 19) +
                C4[5] * xz * (7.0f * zz - 3.0f) * rest(ch, 20) +
                C4[6] * (xx - yy) * (7.0f * zz - 1.0f) * rest(ch, 21) +
                C4[7] * xz * (xx - 3.0f * yy) * rest(ch, 22) +
                C4[8] * (xx * (xx - 3.0f * yy) - yy * (3.0f * xx - yy)) * rest(ch, 23);
          }
        }
      }
      res[ch] = r;
    }
  }
  for (int ch = 0; ch < 3; ++ch) {
    float rgb = std::min(std::max(res[ch] + 0.5f, 0.0f), 1.0f);
    out_rgba[ch] = rgb * alpha;
  }
  out_rgba[3] = alpha;
}

bool covered(float alpha, float r2) {
  // fs_main discards, wgsl:358-367.
  if (r2 > 1.0f) return false;
  float g = std::exp(-4.5f * r2);
  return alpha * g > (1.0f / 256.0f);
}

// Shared r^2 definition (the HIP raster evaluates the same expression):
// local = A^{-1} (p_ndc - center), A = [axis_u axis_v] (wgsl:346-347, 358),
// with the pixel centre p_ndc = (2(x+.5)/W - 1, 1 - 2(y+.5)/H).
struct InvA {
  float i00, i01, i10, i11;
};
InvA inverse_axes(const Proj &p) {
  float det = p.ux * p.vy - p.vx * p.uy;
  return InvA{p.vy / det, -p.vx / det, -p.uy / det, p.ux / det};
}
float pixel_ndc_x(uint32_t x, uint32_t w) { return (2.0f * ((float)x + 0.5f)) / (float)w - 1.0f; }
float pixel_ndc_y(uint32_t y, uint32_t h) { return 1.0f - (2.0f * ((float)y + 0.5f)) / (float)h; }
float local_r2(const InvA &ia, float cx, float cy, float pnx, float pny) {
  float dx = pnx - cx;
  float dy = pny - cy;
  float lx = std::fma(ia.i01, dy, ia.i00 * dx);
  float ly = std::fma(ia.i11, dy, ia.i10 * dx);
  return std::fma(ly, ly, lx * lx);
}

// Tile rectangle of the covered set {r2 <= cut} of one splat: the set is the
// ellipse center + A*l, |l| <= sqrt(cut), whose NDC half extents are
// sqrt(cut)*|row of A|; one pixel of margin on each side.
bool tile_rect(const Proj &p, float cut, uint32_t w, uint32_t h, uint32_t tile, int rect[4]) {
  if (!p.valid || !(cut >= 0.0f)) return false;
  float sc = std::sqrt(cut);
  float ex = sc * std::sqrt(p.ux * p.ux + p.vx * p.vx);
  float ey = sc * std::sqrt(p.uy * p.uy + p.vy * p.vy);
#if ORACLE_MUTATION == 7
  ex = sc * (std::fabs(p.ux) + std::fabs(p.vx));
  ey = sc * (std::fabs(p.uy) + std::fabs(p.vy));
#endif
  float hw = 0.5f * (float)w, hh = 0.5f * (float)h;
  float cxp = (p.cx + 1.0f) * hw;
  float cyp = (1.0f - p.cy) * hh;
  float exw = ex * hw + 1.0f;
  float eyh = ey * hh + 1.0f;
  float x0f = std::ceil(cxp - exw - 0.5f), x1f = std::floor(cxp + exw - 0.5f);
  float y0f = std::ceil(cyp - eyh - 0.5f), y1f = std::floor(cyp + eyh - 0.5f);
  if (x1f < 0.0f || y1f < 0.0f || x0f > (float)(w - 1) || y0f > (float)(h - 1) || x0f > x1f ||
      y0f > y1f)
    return false;
  int x0 = (int)std::max(x0f, 0.0f), x1 = (int)std::min(x1f, (float)(w - 1));
  int y0 = (int)std::max(y0f, 0.0f), y1 = (int)std::min(y1f, (float)(h - 1));
  rect[0] = x0 / (int)tile;
  rect[1] = y0 / (int)tile;
  rect[2] = x1 / (int)tile;
  rect[3] = y1 / (int)tile;
  return true;
}

double g_timings[4] = {0, 0, 0, 0};

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" {

// precompute_alpha_values / sigmoid (lib.rs:1251-1261) and
// precompute_world_covariances (lib.rs:1654-1674).
void oracle_precompute(const OracleScene *s, float *alpha_out, float *cov6_out) {
  for (uint32_t i = 0; i < s->count; ++i) {
    float o = s->opacity[i];
    float sig = 1.0f / (1.0f + std::exp(-o));
    alpha_out[i] = std::min(std::max(sig, 0.0f), 1.0f);

    const float *sc = s->scale_xyz + 3 * (size_t)i;
    float sx = std::max(std::exp(sc[0]), 1e-6f);
    float sy = std::max(std::exp(sc[1]), 1e-6f);
    float sz = std::max(std::exp(sc[2]), 1e-6f);
    float obj[3][3] = {{sx * sx, 0.0f, 0.0f}, {0.0f, sy * sy, 0.0f}, {0.0f, 0.0f, sz * sz}};
    const float *q = s->rotation_xyzw + 4 * (size_t)i;
    float rot[3][3];
    quat_to_mat3(quat_normalize(Q{q[0], q[1], q[2], q[3]}), rot);
    float rt[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) rt[r][c] = rot[c][r];
    float tmp[3][3], world[3][3];
    mat3_mul(rot, obj, tmp);
    mat3_mul(tmp, rt, world);
    float *out = cov6_out + 6 * (size_t)i;
    out[0] = world[0][0];  // CameraCovarianceTerms::from_matrix, lib.rs:1518-1527
    out[1] = world[0][1];
    out[2] = world[0][2];
    out[3] = world[1][1];
    out[4] = world[1][2];
    out[5] = world[2][2];
  }
}

void oracle_view_rows(const float rot[4], float rows_out[9]) {
  float m[3][3];
  quat_to_mat3(quat_inverse(Q{rot[0], rot[1], rot[2], rot[3]}), m);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) rows_out[3 * r + c] = m[r][c];
}

int oracle_camera_valid(const OracleCamera *c) {
  for (int i = 0; i < 3; ++i)
    if (!std::isfinite(c->position[i])) return 0;
  for (int i = 0; i < 4; ++i)
    if (!std::isfinite(c->rotation_xyzw[i])) return 0;
  float norm2 = 0.0f;  // Iterator::sum over the 4 squares, left to right
  for (int i = 0; i < 4; ++i) norm2 += c->rotation_xyzw[i] * c->rotation_xyzw[i];
  if (norm2 <= 0.0f || !std::isfinite(norm2)) return 0;
  if (std::fabs(norm2 - 1.0f) > 1.0e-3f) return 0;
  if (!std::isfinite(c->vertical_fov_radians) || !std::isfinite(c->near_plane) ||
      !std::isfinite(c->far_plane))
    return 0;
  if (c->vertical_fov_radians <= 0.0f || c->vertical_fov_radians >= 3.14159265358979323846f ||
      c->near_plane <= 0.0f || c->far_plane <= c->near_plane)
    return 0;
  return 1;
}

// preprocess_visible_into, lib.rs:1263-1296; is_visible :969-971;
// depth_to_key :973-976.
uint32_t oracle_preprocess(const OracleScene *s, const OracleCamera *cam, uint32_t *keys_out,
                           uint32_t *indices_out) {
  View v = make_view(cam);
  uint32_t n = 0;
  for (uint32_t i = 0; i < s->count; ++i) {
    float depth = depth_of(v, s->positions + 3 * (size_t)i);
    if (depth >= cam->near_plane && depth <= cam->far_plane) {
      indices_out[n] = i;
      keys_out[n] = f32_bits(std::max(depth, 0.0f));
      ++n;
    }
  }
  return n;
}

void oracle_sort_values_by_keys(const uint32_t *keys, uint32_t *values, uint64_t len) {
  if (len <= 1) return;  // sort/src/lib.rs:45-48
  std::vector<uint64_t> packed(len);
  for (uint64_t i = 0; i < len; ++i)
    packed[i] = ((uint64_t)keys[i] << 32) | (uint64_t)(uint32_t)(~values[i]);
  radix_sort_key_bits(packed);
  for (uint64_t i = 0; i < len; ++i) values[i] = ~(uint32_t)packed[i];  // unpack_values
}

void oracle_project(const OracleScene *s, const OracleCamera *cam, uint32_t width, uint32_t height,
                    float *out12) {
  View v = make_view(cam);
  FrameConsts k = frame_consts(v, width, height);
  std::vector<float> alpha(s->count), cov(6 * (size_t)s->count);
  oracle_precompute(s, alpha.data(), cov.data());
  for (uint32_t i = 0; i < s->count; ++i) {
    Proj p = project_splat(v, k, s->positions + 3 * (size_t)i, cov.data() + 6 * (size_t)i, alpha[i]);
    float *o = out12 + 12 * (size_t)i;
    std::fill(o, o + 12, 0.0f);
    if (!p.valid) continue;
    float rgba[4];
    eval_color(s, i, v, p.alpha, rgba);
    o[0] = 1.0f;
    o[1] = p.cx;
    o[2] = p.cy;
    o[3] = p.ux;
    o[4] = p.uy;
    o[5] = p.vx;
    o[6] = p.vy;
    o[7] = rgba[0];
    o[8] = rgba[1];
    o[9] = rgba[2];
    o[10] = rgba[3];
  }
}

float oracle_coverage_cut(float alpha) {
  if (!covered(alpha, 0.0f)) return -1.0f;
  uint32_t lo = 0, hi = f32_bits(1.0f);  // covered(lo) holds
  if (covered(alpha, 1.0f)) return 1.0f;
  while (hi - lo > 1) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (covered(alpha, bits_f32(mid)))
      lo = mid;
    else
      hi = mid;
  }
  return bits_f32(lo);
}

void oracle_tile_counts(uint32_t count, const float *proj12, uint32_t width, uint32_t height,
                        uint32_t tile, uint32_t *counts_out) {
  uint32_t tx = (width + tile - 1) / tile, ty = (height + tile - 1) / tile;
  std::fill(counts_out, counts_out + (size_t)tx * ty, 0u);
  for (uint32_t i = 0; i < count; ++i) {
    const float *o = proj12 + 12 * (size_t)i;
    Proj p{o[0] != 0.0f, o[1], o[2], o[3], o[4], o[5], o[6], o[10]};
    float cut = std::min(oracle_coverage_cut(p.alpha), 1.0f);
    int r[4];
    if (!tile_rect(p, cut, width, height, tile, r)) continue;
    for (int y = r[1]; y <= r[3]; ++y)
      for (int x = r[0]; x <= r[2]; ++x) counts_out[(size_t)y * tx + x] += 1;
  }
}

uint32_t oracle_render(const OracleScene *s, const OracleCamera *cam, uint32_t width,
                       uint32_t height, int threads, float *rgba) {
  return oracle_render_window(s, cam, width, height, 0, 0, width, height, threads, rgba);
}

}  // extern "C"

namespace {

// A frame (window [wx0, wx1) x [wy0, wy1)) from the load-time values alpha
// and cov6 of oracle_precompute (Renderer::load_scene computes them once,
// lib.rs:499-527). The preprocess and the sort run on one thread, as the
// reference's do; the projection of the sorted instances runs on `threads`
// (the reference's rayon par_iter, lib.rs:1066-1101) and so does the blend.
uint32_t render_window_prepared(const OracleScene *s, const float *alpha, const float *cov6,
                                const OracleCamera *cam, uint32_t width, uint32_t height, uint32_t wx0,
                                uint32_t wy0, uint32_t wx1, uint32_t wy1, int threads, float *rgba) {
  if (!oracle_camera_valid(cam) || width == 0 || height == 0) return UINT32_MAX;
  if (wx0 >= wx1 || wy0 >= wy1 || wx1 > width || wy1 > height) return UINT32_MAX;
  const uint32_t ww = wx1 - wx0, wh = wy1 - wy0;
  int nthreads = 1;
#ifdef _OPENMP
  nthreads = threads > 0 ? threads : omp_get_max_threads();
#else
  (void)threads;
#endif
  auto t0 = std::chrono::steady_clock::now();
  std::vector<uint32_t> keys(s->count), order(s->count);
  uint32_t vis = oracle_preprocess(s, cam, keys.data(), order.data());
  g_timings[0] = ms_since(t0);
  t0 = std::chrono::steady_clock::now();
  oracle_sort_values_by_keys(keys.data(), order.data(), vis);
  g_timings[1] = ms_since(t0);

  // Projection of the sorted instances (vs_main fetches sorted[k], wgsl:341-345).
  t0 = std::chrono::steady_clock::now();
  View v = make_view(cam);
  FrameConsts k = frame_consts(v, width, height);
  struct Inst {
    Proj p;
    InvA ia;
    float rgba[4];
    int x0, y0, x1, y1;
  };
  // Instances in sorted order; one that is not drawn or reaches no pixel of
  // the window keeps x0 > x1 and is dropped afterwards.
  std::vector<Inst> all(vis);
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (uint32_t j = 0; j < vis; ++j) {
    uint32_t i = order[j];
    Inst &in = all[j];
    in.x0 = 1;
    in.x1 = 0;
    Proj p = project_splat(v, k, s->positions + 3 * (size_t)i, cov6 + 6 * (size_t)i, alpha[i]);
    if (!p.valid) continue;
    in.p = p;
    in.ia = inverse_axes(p);
    eval_color(s, i, v, p.alpha, in.rgba);
    // Pixel bounds of the instanced quad center +- s*axis_u +- s*axis_v
    // (vs_main + alpha_extent_scale, wgsl:70-77, 346-350), one pixel of margin.
    float sq = std::sqrt(std::min(std::max(std::log(std::max(p.alpha, 1e-12f) * 256.0f) / 4.5f, 0.0f), 1.0f));
    float ex = sq * (std::fabs(p.ux) + std::fabs(p.vx));
    float ey = sq * (std::fabs(p.uy) + std::fabs(p.vy));
    float x0 = ((p.cx - ex) + 1.0f) * 0.5f * (float)width - 1.5f;
    float x1 = ((p.cx + ex) + 1.0f) * 0.5f * (float)width + 0.5f;
    float y0 = (1.0f - (p.cy + ey)) * 0.5f * (float)height - 1.5f;
    float y1 = (1.0f - (p.cy - ey)) * 0.5f * (float)height + 0.5f;
    in.x0 = (int)std::max(std::floor(x0), 0.0f);
    in.x1 = (int)std::min(std::ceil(x1), (float)width - 1.0f);
    in.y0 = (int)std::max(std::floor(y0), 0.0f);
    in.y1 = (int)std::min(std::ceil(y1), (float)height - 1.0f);
    // clipped to the window (a pixel's result depends only on the splats
    // that reach it, in the same order)
    in.x0 = std::max(in.x0, (int)wx0);
    in.x1 = std::min(in.x1, (int)wx1 - 1);
    in.y0 = std::max(in.y0, (int)wy0);
    in.y1 = std::min(in.y1, (int)wy1 - 1);
    if (in.y0 > in.y1) in.x1 = in.x0 - 1;
  }
  std::vector<Inst> inst;
  inst.reserve(vis);
  for (const Inst &in : all)
    if (in.x0 <= in.x1) inst.push_back(in);
  g_timings[2] = ms_since(t0);

  // Clear to transparent black (lib.rs:2746), then premultiplied "over"
  // back to front: dst = src + dst*(1 - src.a) (draw_pass.rs:75).
  t0 = std::chrono::steady_clock::now();
  std::fill(rgba, rgba + (size_t)ww * wh * 4, 0.0f);
  int bands = nthreads * 4;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int band = 0; band < bands; ++band) {
    int by0 = (int)wy0 + (int)((int64_t)wh * band / bands);
    int by1 = (int)wy0 + (int)((int64_t)wh * (band + 1) / bands) - 1;
    if (by1 < by0) continue;
    for (const Inst &in : inst) {
      int y0 = std::max(in.y0, by0), y1 = std::min(in.y1, by1);
      for (int y = y0; y <= y1; ++y) {
        float pny = pixel_ndc_y((uint32_t)y, height);
        for (int x = in.x0; x <= in.x1; ++x) {
          float pnx = pixel_ndc_x((uint32_t)x, width);
          float r2 = local_r2(in.ia, in.p.cx, in.p.cy, pnx, pny);
          if (r2 > 1.0f) continue;
          float g = std::exp(-4.5f * r2);
          float a = in.rgba[3] * g;
          if (a <= 1.0f / 256.0f) continue;
          float *d = rgba + ((size_t)(y - (int)wy0) * ww + (size_t)(x - (int)wx0)) * 4;
          float one_minus = 1.0f - a;
          d[0] = in.rgba[0] * g + d[0] * one_minus;
          d[1] = in.rgba[1] * g + d[1] * one_minus;
          d[2] = in.rgba[2] * g + d[2] * one_minus;
          d[3] = a + d[3] * one_minus;
        }
      }
    }
  }
  g_timings[3] = ms_since(t0);
  return vis;
}

}  // namespace

extern "C" {

// The same frame restricted to pixels [wx0, wx1) x [wy0, wy1): the window is
// written row-major into rgba ((wx1-wx0) * (wy1-wy0) * 4 floats), every
// pixel computed exactly as the full render computes it.
uint32_t oracle_render_window(const OracleScene *s, const OracleCamera *cam, uint32_t width, uint32_t height,
                              uint32_t wx0, uint32_t wy0, uint32_t wx1, uint32_t wy1, int threads, float *rgba) {
  std::vector<float> alpha(s->count), cov(6 * (size_t)s->count);
  oracle_precompute(s, alpha.data(), cov.data());
  return render_window_prepared(s, alpha.data(), cov.data(), cam, width, height, wx0, wy0, wx1, wy1, threads,
                                rgba);
}

// oracle_render with the load-time values of oracle_precompute passed in.
uint32_t oracle_render_prepared(const OracleScene *s, const float *alpha, const float *cov6,
                                const OracleCamera *cam, uint32_t width, uint32_t height, int threads,
                                float *rgba) {
  return render_window_prepared(s, alpha, cov6, cam, width, height, 0, 0, width, height, threads, rgba);
}

// Front-to-back statement of the blend's work, for the raster roofline: per
// pixel, the visible drawn splats nearest first; an evaluation passes when
// r2 <= cut (the splat's coverage bound, oracle_coverage_cut capped at 1);
// each passing one updates T *= 1 - alpha*expf(-4.5*r2) and the pixel stops
// once T < 1e-4. Returns the number of passing evaluations.
uint64_t oracle_blend_evals(const OracleScene *s, const OracleCamera *cam, uint32_t width, uint32_t height) {
  if (!oracle_camera_valid(cam) || width == 0 || height == 0) return 0;
  std::vector<uint32_t> keys(s->count), order(s->count);
  uint32_t vis = oracle_preprocess(s, cam, keys.data(), order.data());
  oracle_sort_values_by_keys(keys.data(), order.data(), vis);
  View v = make_view(cam);
  FrameConsts k = frame_consts(v, width, height);
  std::vector<float> alpha(s->count), cov(6 * (size_t)s->count);
  oracle_precompute(s, alpha.data(), cov.data());
  struct Inst {
    Proj p;
    InvA ia;
    float cut;
    int x0, y0, x1, y1;
  };
  std::vector<Inst> inst;
  for (uint32_t j = vis; j-- > 0;) {  // front to back
    uint32_t i = order[j];
    Proj p = project_splat(v, k, s->positions + 3 * (size_t)i, cov.data() + 6 * (size_t)i, alpha[i]);
    if (!p.valid) continue;
    Inst in;
    in.p = p;
    in.ia = inverse_axes(p);
    in.cut = std::min(oracle_coverage_cut(p.alpha), 1.0f);
    if (in.cut < 0.0f) continue;
    int r[4];
    if (!tile_rect(p, in.cut, width, height, 1, r)) continue;  // pixel bounds
    in.x0 = r[0];
    in.y0 = r[1];
    in.x1 = r[2];
    in.y1 = r[3];
    inst.push_back(in);
  }
  std::vector<float> T((size_t)width * height, 1.0f);
  uint64_t evals = 0;
  for (const Inst &in : inst) {
    for (int y = in.y0; y <= in.y1; ++y) {
      float pny = pixel_ndc_y((uint32_t)y, height);
      for (int x = in.x0; x <= in.x1; ++x) {
        float &t = T[(size_t)y * width + x];
        if (t < 1e-4f) continue;
        float r2 = local_r2(in.ia, in.p.cx, in.p.cy, pixel_ndc_x((uint32_t)x, width), pny);
        if (!(r2 <= in.cut)) continue;
        ++evals;
        t = t * (1.0f - in.p.alpha * std::exp(-4.5f * r2));
      }
    }
  }
  return evals;
}

void oracle_last_timings(double out_ms[4]) {
  for (int i = 0; i < 4; ++i) out_ms[i] = g_timings[i];
}

}  // extern "C"
