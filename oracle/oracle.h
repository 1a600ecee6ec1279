/*
 * oracle.h -- CPU restatement of the gsplat-rs SortedAlpha render path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it. The product library (libgsplat_ffi_c.so) never links, calls or
 * falls back to it.
 *
 * Every function restates reference code at the cited file:line (paths are
 * relative to the gsplat-rs repository root). Floating point follows the
 * reference's evaluation order with contraction off (Rust never fuses), so
 * depth keys and projected quantities are bit-exact statements of the
 * reference arithmetic on IEEE hardware. Where the reference's arithmetic is
 * done by a GPU driver with no canonical operation order (the rasteriser's
 * interpolation of `local`), the order used here is stated in the function's
 * comment and is the definition the HIP kernel must match.
 *
 * Pinning (what checks each function; details in tests/test_oracle_*.py):
 *   oracle_precompute        closed forms (identity quat -> diag(e^{2s}),
 *                            90-degree rotation swaps axes), sigmoid values
 *   oracle_view_rows         quat_inverse([0,0,0,2]) = identity
 *                            (render lib.rs:4113-4116), orthonormality
 *   oracle_preprocess        order/visibility KAT [1,2,0] (lib.rs:4062-4089),
 *                            brute force
 *   oracle_sort_values_by_keys  KATs of sort/src/lib.rs:602-688 and
 *                            std::stable_sort brute force over LCG inputs
 *   oracle_project           invariants of lib.rs:4442-4528, closed-form
 *                            isotropic splat, Legendre-based SH basis check
 *   oracle_render            conformance means of
 *                            tests/conformance_sorted_alpha.rs:43-76,
 *                            analytic single-splat image, brute force
 *   oracle_tile_counts       brute force over the oracle's own coverage
 *                            (every covered pixel lies in a counted tile)
 *   oracle_blend_evals       brute force per pixel over all splats on tiny
 *                            scenes; single opaque splat closed form
 */
#ifndef GSPLAT_ORACLE_H
#define GSPLAT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same layout as GsplatCamera (gsplat.h:86-92). */
typedef struct OracleCamera {
  float position[3];
  float rotation_xyzw[4];
  float vertical_fov_radians;
  float near_plane;
  float far_plane;
} OracleCamera;

/* Scene in the reference's SceneBuffers layout (core/src/lib.rs:201-213). */
typedef struct OracleScene {
  uint32_t count;
  uint32_t sh_degree;
  const float *positions;     /* 3N */
  const float *opacity;       /* N  */
  const float *scale_xyz;     /* 3N */
  const float *rotation_xyzw; /* 4N */
  const float *color_dc;      /* 3N */
  const float *sh_rest;       /* N*3*((D+1)^2-1), channel-major per splat, or NULL */
} OracleScene;

/* alpha = clamp(sigmoid(o),0,1); cov6 = (xx,xy,xz,yy,yz,zz) of R diag(s^2) R^T. */
void oracle_precompute(const OracleScene *scene, float *alpha_out, float *cov6_out);

/* View rotation rows of quat_to_mat3(quat_inverse(q)); row-major 9 floats. */
void oracle_view_rows(const float rotation_xyzw[4], float rows_out[9]);

/* Camera::validate (core/src/lib.rs:39-107): 1 if valid. */
int oracle_camera_valid(const OracleCamera *camera);

/* preprocess_visible_into: visible indices ascending + depth keys; returns V. */
uint32_t oracle_preprocess(const OracleScene *scene, const OracleCamera *camera,
                           uint32_t *keys_out, uint32_t *indices_out);

/* CpuSortBackend::sort_values_by_keys (stable, descending, 8-bit LSD). */
void oracle_sort_values_by_keys(const uint32_t *keys, uint32_t *values, uint64_t len);

/* project_splat + eval_color per splat, 12 floats per splat:
 * [drawn, cx, cy, ux, uy, vx, vy, r*a, g*a, b*a, a, 0]. */
void oracle_project(const OracleScene *scene, const OracleCamera *camera, uint32_t width,
                    uint32_t height, float *out12);

/* Largest r2 in [0,1] with r2 <= 1 && alpha*expf(-4.5*r2) > 1/256, or -1. */
float oracle_coverage_cut(float alpha);

/* (tile, splat) pair counts per tile_size x tile_size tile of the drawn
 * splats in `proj12` (oracle_project layout) with coverage cuts from alpha. */
void oracle_tile_counts(uint32_t count, const float *proj12, uint32_t width, uint32_t height,
                        uint32_t tile_size, uint32_t *counts_out);

/* Full frame: preprocess, sort, project, back-to-front premultiplied "over"
 * into a cleared fp32 RGBA target (width*height*4 floats, rows top first).
 * Returns V (visible count) or UINT32_MAX on an invalid camera/config.
 * `threads` = OpenMP threads for the projection of the sorted instances and
 * the raster (0 = default); the preprocess and the sort take one thread, as
 * the reference's do. */
uint32_t oracle_render(const OracleScene *scene, const OracleCamera *camera, uint32_t width,
                       uint32_t height, int threads, float *rgba_out);
/* oracle_render from the load-time values of oracle_precompute (alpha N,
 * cov6 6N), which the reference computes once per scene load rather than per
 * frame (render-wgpu/src/lib.rs:499-527). */
uint32_t oracle_render_prepared(const OracleScene *scene, const float *alpha, const float *cov6,
                                const OracleCamera *camera, uint32_t width, uint32_t height, int threads,
                                float *rgba_out);
/* oracle_render restricted to the pixel window [wx0, wx1) x [wy0, wy1),
 * written row-major into a (wx1-wx0) x (wy1-wy0) RGBA float buffer; each
 * pixel equals the full render's. UINT32_MAX for an empty or out-of-range
 * window. */
uint32_t oracle_render_window(const OracleScene *scene, const OracleCamera *camera, uint32_t width,
                              uint32_t height, uint32_t wx0, uint32_t wy0, uint32_t wx1, uint32_t wy1, int threads,
                              float *rgba);

/* Number of (pixel, splat) evaluations a front-to-back blend makes that pass
 * the coverage test before each pixel's T < 1e-4 stop (the raster's
 * algorithmic work); 0 on an invalid camera/config. */
uint64_t oracle_blend_evals(const OracleScene *scene, const OracleCamera *camera, uint32_t width,
                            uint32_t height);

/* Stage timers of the last oracle_render in ms: preprocess, sort, project, raster. */
void oracle_last_timings(double out_ms[4]);

#ifdef __cplusplus
}
#endif

#endif
